"""Inviscid face-flux schemes (reference `wmles/schemes.py`).

The objects describe the scheme; the face fluxes themselves are evaluated inside
the fused HIP flux/divergence kernel (`csrc/rhs.hip`, `csrc/physics.h`).
"""
from . import _lib


class _Scheme:
    def __call__(self, blk, geom, w, gas, d, theta=None):
        raise NotImplementedError("face fluxes are evaluated inside the HIP step; use Solver.rhs")


class KEP(_Scheme):
    """Kinetic-energy-preserving split form, order 2 or 4 (schemes.py:57-70)."""
    code = _lib.SCHEME_KEP

    def __init__(self, order=4):
        assert order in (2, 4)
        self.order = order


class KEEP(_Scheme):
    """Kinetic-energy and entropy preserving split form, order 2 or 4 (schemes.py:72-98)."""
    code = _lib.SCHEME_KEEP

    def __init__(self, order=4):
        assert order in (2, 4)
        self.order = order


class MUSCLHLLC(_Scheme):
    """MUSCL (van Albada / minmod) reconstruction of primitives + HLLC (schemes.py:140-152)."""
    code = _lib.SCHEME_MUSCLHLLC

    def __init__(self, limiter="albada"):
        if limiter not in ("albada", "minmod"):
            raise KeyError(limiter)
        self.limiter = limiter


class Hybrid(_Scheme):
    """F = (1-s) F_central + s F_upwind, s = ramp(max(theta_L, theta_R)) (schemes.py:155-166)."""
    code = _lib.SCHEME_HYBRID

    def __init__(self, central=None, upwind=None, th0=0.6, th1=0.9):
        self.c = central or KEP(4)
        self.u = upwind or MUSCLHLLC()
        if not isinstance(self.c, (KEP, KEEP)) or not isinstance(self.u, MUSCLHLLC):
            raise NotImplementedError("Hybrid supports a KEP/KEEP central and a MUSCLHLLC upwind part")
        self.th0, self.th1 = th0, th1

"""Inviscid face-flux schemes (reference `wmles/schemes.py`).

The objects describe the scheme to the fused HIP step (`csrc/flux.hip`), and keep the
reference's callable protocol `scheme(blk, geom, w, gas, d, theta) -> F`
(`wmles/schemes.py:3-5`): called directly, the flux goes through `wmx_scheme_flux`,
which evaluates the same FP64 device code (`csrc/physics.h`) as the step.
"""
import ctypes as C

import torch

from . import _lib
from .backend import asarray


class _Scheme:
    def __call__(self, blk, geom, w, gas, d, theta=None):
        """Inviscid flux incl. area through the n_d + 1 interior faces of direction d,
        (5, .. n_d + 1 .., n_t1, n_t2), from the ghosted primitives w = [rho, u, v, w, p]
        and (Hybrid) the Ducros sensor theta."""
        from .solver import numerics_desc
        if d not in (0, 1, 2):
            raise ValueError("d must be 0, 1 or 2")
        w = asarray(w)
        if tuple(w.shape) != (5,) + tuple(blk.N):
            raise ValueError(f"w must be (5,)+{tuple(blk.N)}, got {tuple(w.shape)}")
        S = asarray(geom.S[d])
        th = None
        if theta is not None:
            th = asarray(theta)
            if tuple(th.shape) != tuple(blk.N):
                raise ValueError(f"theta must be {tuple(blk.N)}, got {tuple(th.shape)}")
        if isinstance(self, Hybrid) and th is None:
            raise TypeError("Hybrid needs the shock sensor theta")
        nd = numerics_desc(gas, self, None, 0.8, 2, (), ())
        n = (C.c_int32 * 3)(*blk.n)
        shape = (5,) + tuple(blk.n[a] + (a == d) for a in range(3))
        F = torch.empty(shape, dtype=torch.float64, device=w.device)
        _lib.check(_lib.lib.wmx_scheme_flux(C.byref(nd), n, int(d), C.c_void_p(w.data_ptr()),
                                            C.c_void_p(S.data_ptr()),
                                            C.c_void_p(th.data_ptr()) if th is not None else None,
                                            C.c_void_p(F.data_ptr()), _lib.stream_ptr()))
        return F


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

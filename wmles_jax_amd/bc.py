"""Boundary conditions (reference `wmles/bc.py`).

Each BC is a descriptor the HIP ghost sweep executes; the class names, constructor
arguments and the `fill` / `fill_at_time` / `flux_override` / `wall_flux` protocol
follow the reference (`bc.py:1-8,28-35`).
"""
import numpy as np
import torch

from . import _lib


class BC:
    """Base class (bc.py:28-35)."""
    flux_override = False
    kind = None

    def _fill_desc(self, fd, ctx=None):
        raise NotImplementedError(
            f"{type(self).__name__}: boundary conditions without a HIP descriptor are not "
            "supported by the HIP step")

    def fill(self, q, blk, geom, gas, axis, side):
        """Return q with the ghost layers of face (axis, side) filled (one face of one
        block, executed by the HIP fill kernel)."""
        return self.fill_at_time(q, blk, geom, gas, axis, side, 0.0)

    def fill_at_time(self, q, blk, geom, gas, axis, side, time):
        from .mesh import block_desc
        from .solver import numerics_desc
        out = q.clone().contiguous()
        saved = dict(blk.faces)
        try:
            blk.faces = {(a, s): SlipWall() for a in range(3) for s in range(2)}
            blk.faces[(axis, side)] = self
            d = block_desc(blk, None, None, geom)
        finally:
            blk.faces = saved
        nd = numerics_desc(gas, None, None, 0.8, 2, (), ())
        _lib.check(_lib.lib.wmx_fill_face(_lib.C.byref(d), axis, side, _lib.C.byref(nd),
                                          out.data_ptr(), float(time), _lib.stream_ptr()))
        return out


class SlipWall(BC):
    """Symmetry / inviscid wall: mirror with the normal velocity reflected (bc.py:38-47)."""
    kind = _lib.BC_SLIP

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_SLIP


class NoSlipWall(BC):
    """u_g = -u; T_wall=None -> adiabatic (bc.py:50-64)."""
    kind = _lib.BC_NOSLIP

    def __init__(self, T_wall=None):
        self.T_wall = T_wall

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_NOSLIP
        fd.T_wall = -1.0 if self.T_wall is None else float(self.T_wall)


class Dirichlet(BC):
    """Fixed primitive state [rho, u, v, w, p] (bc.py:67-77)."""
    kind = _lib.BC_DIRICHLET

    def __init__(self, prim):
        self.prim = np.asarray(prim, float)

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_DIRICHLET
        for i in range(5):
            fd.prim[i] = float(self.prim[i])


class CharacteristicFarfield(BC):
    """Linearised characteristic far field for primitive [rho, u, v, w, p] (bc.py:80-115)."""
    kind = _lib.BC_FARFIELD

    def __init__(self, prim):
        self.prim = np.asarray(prim, float)

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_FARFIELD
        for i in range(5):
            fd.prim[i] = float(self.prim[i])


class Extrapolate(BC):
    """Zero gradient (bc.py:118-125)."""
    kind = _lib.BC_EXTRAPOLATE

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_EXTRAPOLATE


class WallModel(BC):
    """Wall-modelled wall: slip ghosts + wall-face flux [0, p n + tau_w t, -q_w] |S|
    (bc.py:128-166).  T_wall=None -> adiabatic."""
    flux_override = True
    kind = _lib.BC_WALLMODEL

    def __init__(self, model, T_wall=None, n_match=1):
        self.model, self.T_wall, self.n_match = model, T_wall, n_match
        self._slip = SlipWall()

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_WALLMODEL
        fd.T_wall = -1.0 if self.T_wall is None else float(self.T_wall)
        fd.n_match = int(self.n_match)

    def fill(self, q, blk, geom, gas, axis, side):
        return self._slip.fill(q, blk, geom, gas, axis, side)

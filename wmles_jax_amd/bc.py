"""Boundary conditions (reference `wmles/bc.py`).

Each BC is a descriptor the HIP ghost sweep executes; the class names, constructor
arguments and the `fill` / `fill_at_time` / `flux_override` / `wall_flux` protocol
follow the reference (`bc.py:1-8,28-35`).
"""
import numpy as np
import torch

from . import _lib


def is_external(bc):
    """A BC with no HIP descriptor of its own (a user subclass of BC implementing `fill` /
    `fill_at_time` in Python): the solver runs its fill itself, in the sweep's order."""
    return isinstance(bc, BC) and type(bc)._fill_desc is BC._fill_desc


class BC:
    """Base class (bc.py:28-35).  Subclasses written by users implement
    `fill(q, blk, geom, gas, axis, side) -> q` (and optionally `fill_at_time`) on torch
    tensors or arrays; the solver calls them between the HIP sweep phases."""
    flux_override = False
    kind = None

    def _fill_desc(self, fd, ctx=None):
        if getattr(self, "flux_override", False):
            raise NotImplementedError(f"{type(self).__name__}: a user BC that overrides the wall flux "
                                      "has no HIP step path")
        fd.kind = _lib.BC_EXTERNAL

    def fill(self, q, blk, geom, gas, axis, side):
        """Return q with the ghost layers of face (axis, side) filled (one face of one
        block, executed by the HIP fill kernel)."""
        return self._hip_fill(q, blk, geom, gas, axis, side, 0.0)

    def fill_at_time(self, q, blk, geom, gas, axis, side, time):
        """The fill at a stage time (bc.py:34-35: the time-independent fill by default)."""
        if type(self).fill is not BC.fill:
            return self.fill(q, blk, geom, gas, axis, side)
        return self._hip_fill(q, blk, geom, gas, axis, side, time)

    def _hip_fill(self, q, blk, geom, gas, axis, side, time):
        if is_external(self):
            raise NotImplementedError(f"{type(self).__name__} implements neither fill nor a HIP descriptor")
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

    def wall_flux(self, w, T, blk, geom, gas, axis, side):
        """(F, tau_w, q_w) on face (axis, side): F = +-[0, p_b n_out + tau_w t, -q_w] |S|
        (5, n_t1, n_t2) with the sign of the face's outward side (bc.py:145-166).  The
        equilibrium model runs in the HIP kernel `wmx_wall_flux`; any other model
        callable model(gas, h, U, T_h, p_h, T_wall) gets the same sampled inputs."""
        from .backend import asarray
        from .wall_model import EquilibriumWallModel
        w, T = asarray(w), asarray(T)
        if tuple(w.shape) != (5,) + tuple(blk.N) or tuple(T.shape) != tuple(blk.N):
            raise ValueError(f"w must be (5,)+{tuple(blk.N)} and T {tuple(blk.N)}")
        if type(self.model) is EquilibriumWallModel:
            return self._wall_flux_hip(w, T, blk, geom, gas, axis, side)
        return self._wall_flux_callable(w, T, blk, geom, gas, axis, side)

    def _wall_flux_hip(self, w, T, blk, geom, gas, axis, side):
        from .mesh import block_desc
        from .solver import numerics_desc
        saved = dict(blk.faces)
        try:
            blk.faces = {(a, s): SlipWall() for a in range(3) for s in range(2)}
            blk.faces[(axis, side)] = self
            d = block_desc(blk, None, None, geom)
        finally:
            blk.faces = saved
        nd = numerics_desc(gas, None, None, 0.8, 2, (), (), wall_model=self.model)
        t = [blk.n[a] for a in range(3) if a != axis]
        F = torch.empty((5,) + tuple(t), dtype=torch.float64, device=w.device)
        tau = torch.empty(tuple(t), dtype=torch.float64, device=w.device)
        qw = torch.empty_like(tau)
        _lib.check(_lib.lib.wmx_wall_flux(_lib.C.byref(d), int(axis), int(side), _lib.C.byref(nd),
                                          w.data_ptr(), T.data_ptr(), F.data_ptr(), tau.data_ptr(),
                                          qw.data_ptr(), _lib.stream_ptr()))
        return F, tau, qw

    def _wall_flux_callable(self, w, T, blk, geom, gas, axis, side):
        ng, n = blk.ng, blk.n[axis]
        fb = ng if side == 0 else ng + n
        mi = ng + self.n_match - 1 if side == 0 else ng + n - self.n_match
        ci = ng if side == 0 else ng + n - 1
        S = blk.take(geom.S[axis], axis, fb)
        Sm = torch.sqrt((S[0] * S[0] + S[1] * S[1]) + S[2] * S[2])
        n_out = S / Sm if side == 1 else -S / Sm
        d = blk.take(geom.xc, axis, mi) - blk.take(geom.xf[axis], axis, fb)
        h = ((d[0] * -n_out[0] + d[1] * -n_out[1]) + d[2] * -n_out[2])
        u = blk.take(w[1:4], axis, mi)
        un = (u[0] * n_out[0] + u[1] * n_out[1]) + u[2] * n_out[2]
        up = u - un[None] * n_out
        U = torch.sqrt(((up[0] * up[0] + up[1] * up[1]) + up[2] * up[2]) + 1e-20)
        tau, qw = self.model(gas, h, U, blk.take(T, axis, mi), blk.take(w[4], axis, mi), self.T_wall)
        tau, qw = torch.as_tensor(tau, device=w.device), torch.as_tensor(qw, device=w.device)
        pb = blk.take(w[4], axis, ci)
        F = torch.cat([torch.zeros_like(tau)[None], (pb[None] * n_out + tau[None] * (up / U[None])) * Sm[None],
                       (-qw * Sm)[None]], dim=0)
        return (F if side == 1 else -F), tau, qw

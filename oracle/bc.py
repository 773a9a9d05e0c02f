"""Boundary conditions (oracle).  Follows reference `wmles/bc.py:14-166`.

Every fill returns a new state with the ng ghost layers of one face written over
the full tangential extent (ghosts included).
"""
import numpy as np

from .gas import _sum3
from .mesh import sl


def _fill_mirror(q, blk, geom, axis, side, fn):
    """Ghost layer m <- fn(cell layer m from the face, unit +axis normal of the
    boundary face, m) (bc.py:14-25)."""
    ng, n = blk.ng, blk.n[axis]
    fb = ng if side == 0 else ng + n
    S = geom.S[axis][sl(4, axis, slice(fb, fb + 1))]
    nrm = S / np.sqrt(_sum3(S * S))
    q = q.copy()
    for m in range(ng):
        gi = ng - 1 - m if side == 0 else ng + n + m
        ii = ng + m if side == 0 else ng + n - 1 - m
        src = q[sl(q.ndim, axis, slice(ii, ii + 1))]
        q[sl(q.ndim, axis, slice(gi, gi + 1))] = fn(src, nrm, m)
    return q


class BC:
    """bc.py:28-35."""
    flux_override = False

    def fill(self, q, blk, geom, gas, axis, side):
        raise NotImplementedError

    def fill_at_time(self, q, blk, geom, gas, axis, side, time):
        return self.fill(q, blk, geom, gas, axis, side)


class SlipWall(BC):
    """Reflect the normal velocity (bc.py:38-47)."""

    def fill(self, q, blk, geom, gas, axis, side):
        def fn(src, nrm, m):
            w, _ = gas.cons_to_prim(src)
            vel = w[1:4]
            un = _sum3(vel * nrm)
            return gas.prim_to_cons(np.concatenate([w[:1], vel - 2.0 * un * nrm, w[4:5]], 0))
        return _fill_mirror(q, blk, geom, axis, side, fn)


class NoSlipWall(BC):
    """u_g = -u; adiabatic T_g = T or isothermal max(2 T_w - T, 0.1 T_w) (bc.py:50-64)."""

    def __init__(self, T_wall=None):
        self.T_wall = T_wall

    def fill(self, q, blk, geom, gas, axis, side):
        Tw = self.T_wall

        def fn(src, nrm, m):
            w, T = gas.cons_to_prim(src)
            Tg = T if Tw is None else np.maximum(2.0 * Tw - T, 0.1 * Tw)
            rho = w[4] / (gas.R * Tg)
            return gas.prim_to_cons(np.concatenate([rho[None], -w[1:4], w[4:5]], 0))
        return _fill_mirror(q, blk, geom, axis, side, fn)


class Dirichlet(BC):
    """Fixed primitive state (bc.py:67-77)."""

    def __init__(self, prim):
        self.prim = np.asarray(prim, float)

    def fill(self, q, blk, geom, gas, axis, side):
        def fn(src, nrm, m):
            w = self.prim.reshape((5,) + (1,) * (src.ndim - 1)) * np.ones_like(src[:1])
            return gas.prim_to_cons(w)
        return _fill_mirror(q, blk, geom, axis, side, fn)


class CharacteristicFarfield(BC):
    """Linearised characteristic far field (bc.py:80-115)."""

    def __init__(self, prim):
        self.prim = np.asarray(prim, float)

    def fill(self, q, blk, geom, gas, axis, side):
        def fn(src, nrm, m):
            w, T = gas.cons_to_prim(src)
            winf = self.prim.reshape((5,) + (1,) * (src.ndim - 1)) * np.ones_like(src[:1])
            rho_inf, u_inf, p_inf = winf[0], winf[1:4], winf[4]
            c_inf = np.sqrt(gas.gamma * p_inf / rho_inf)
            n_out = nrm if side == 1 else -nrm
            un = _sum3(w[1:4] * n_out)
            un_inf = _sum3(u_inf * n_out)
            c = gas.sound_speed(T)
            drho, dp = w[0] - rho_inf, w[4] - p_inf
            dun = un - un_inf
            a_plus = np.where(un + c > 0.0, dp + rho_inf * c_inf * dun, 0.0)
            a_minus = np.where(un - c > 0.0, dp - rho_inf * c_inf * dun, 0.0)
            dp_bc = 0.5 * (a_plus + a_minus)
            dun_bc = (a_plus - a_minus) / (2.0 * rho_inf * c_inf)
            outgoing = un > 0.0
            entropy = np.where(outgoing, drho - dp / c_inf ** 2, 0.0)
            du_t = w[1:4] - u_inf - dun * n_out
            du_t = np.where(outgoing[None], du_t, 0.0)
            rho_bc = rho_inf + dp_bc / c_inf ** 2 + entropy
            u_bc = u_inf + dun_bc * n_out + du_t
            p_bc = p_inf + dp_bc
            return gas.prim_to_cons(np.concatenate([rho_bc[None], u_bc, p_bc[None]], 0))
        return _fill_mirror(q, blk, geom, axis, side, fn)


class Extrapolate(BC):
    """Zero gradient: every ghost layer = the boundary cell layer (bc.py:118-125)."""

    def fill(self, q, blk, geom, gas, axis, side):
        ng, n = blk.ng, blk.n[axis]
        ii = ng if side == 0 else ng + n - 1
        edge = q[sl(q.ndim, axis, slice(ii, ii + 1))].copy()
        return _fill_mirror(q, blk, geom, axis, side, lambda src, nrm, m: edge)


class WallModel(BC):
    """Slip ghosts + wall-face flux override [0, p_b n + tau t, -q_w] |S| (bc.py:128-166)."""
    flux_override = True

    def __init__(self, model, T_wall=None, n_match=1):
        self.model, self.T_wall, self.n_match = model, T_wall, n_match
        self._slip = SlipWall()

    def fill(self, q, blk, geom, gas, axis, side):
        return self._slip.fill(q, blk, geom, gas, axis, side)

    def wall_flux(self, w, T, blk, geom, gas, axis, side):
        ng, n = blk.ng, blk.n[axis]
        fb = ng if side == 0 else ng + n
        mi = ng + self.n_match - 1 if side == 0 else ng + n - self.n_match
        ci = ng if side == 0 else ng + n - 1
        S = blk.take(geom.S[axis], axis, fb)
        Sm = np.sqrt(_sum3(S * S))
        n_out = S / Sm if side == 1 else -S / Sm
        xf = blk.take(geom.xf[axis], axis, fb)
        xcm = blk.take(geom.xc, axis, mi)
        h = _sum3((xcm - xf) * (-n_out))
        vel = blk.take(w[1:4], axis, mi)
        u_par = vel - _sum3(vel * n_out) * n_out
        U = np.sqrt(_sum3(u_par * u_par) + 1e-20)
        T_h, p_h = blk.take(T, axis, mi), blk.take(w[4], axis, mi)
        p_b = blk.take(w[4], axis, ci)
        tau, qw = self.model(gas, h, U, T_h, p_h, self.T_wall)
        t = u_par / U
        Fout = np.concatenate([np.zeros_like(tau)[None],
                               (p_b * n_out + tau * t) * Sm,
                               (-qw * Sm)[None]], axis=0)
        return (Fout if side == 1 else -Fout), tau, qw

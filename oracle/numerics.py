"""Physics kernels (oracle): fluxes, gradients, SGS, Ducros, wall model.

Follows reference `wmles/schemes.py:15-166`, `wmles/viscous.py:6-51`,
`wmles/sgs.py:6-50`, `wmles/wall_model.py:15-61`.  Pure NumPy float64.
"""
import numpy as np

from .gas import _sum3
from .mesh import sl


# ----------------------------------------------------------------------------- inviscid
def kg_flux(wa, wb, S, gamma):
    """Kennedy-Gruber two-point flux including the area vector (schemes.py:15-23)."""
    rho = 0.5 * (wa[0] + wb[0])
    vel = 0.5 * (wa[1:4] + wb[1:4])
    p = 0.5 * (wa[4] + wb[4])
    ea = 0.5 * (wa[4] / ((gamma - 1.0) * wa[0]) + 0.5 * _sum3(wa[1:4] ** 2)
                + wb[4] / ((gamma - 1.0) * wb[0]) + 0.5 * _sum3(wb[1:4] ** 2))
    un = _sum3(vel * S)
    m = rho * un
    return np.concatenate([m[None], m * vel + p * S, (m * ea + p * un)[None]], axis=0)


def keep2_flux(wa, wb, S, gamma):
    """Kinetic-energy/entropy-preserving two-point flux (schemes.py:25-55)."""
    rho = 0.5 * (wa[0] + wb[0])
    vel = 0.5 * (wa[1:4] + wb[1:4])
    p = 0.5 * (wa[4] + wb[4])
    un = _sum3(vel * S)
    m = rho * un
    K = m * (0.5 * _sum3(wa[1:4] * wb[1:4]))
    ieL = wa[4] / ((gamma - 1.0) * wa[0])
    ieR = wb[4] / ((gamma - 1.0) * wb[0])
    I = m * 0.5 * (ieL + ieR)
    unL = _sum3(wa[1:4] * S)
    unR = _sum3(wb[1:4] * S)
    P = 0.5 * (unR * wa[4] + unL * wb[4])
    return np.concatenate([m[None], m * vel + p * S, (K + I + P)[None]], axis=0)


class _CentralSplit:
    """Order 2: F(L,R); order 4: 2((2/3) F(L,R) - (1/12)(F(L,R2) + F(L2,R)))
    (schemes.py:57-70, 72-98)."""
    two_point = None

    def __init__(self, order=4):
        if order not in (2, 4):
            raise ValueError("order must be 2 or 4")
        self.order = order

    def __call__(self, blk, geom, w, gas, d, theta=None):
        S = blk.slab(geom.S[d], d, 0, 1)
        L, R = blk.slab(w, d, -1, 0), blk.slab(w, d, 0, 1)
        g = gas.gamma
        f2 = type(self).two_point
        F = f2(L, R, S, g)
        if self.order == 2:
            return F
        L2, R2 = blk.slab(w, d, -2, -1), blk.slab(w, d, 1, 2)
        return 2.0 * ((2.0 / 3.0) * F - (1.0 / 12.0) * (f2(L, R2, S, g) + f2(L2, R, S, g)))


class KEP(_CentralSplit):
    two_point = staticmethod(kg_flux)


class KEEP(_CentralSplit):
    two_point = staticmethod(keep2_flux)


def _albada(a, b, eps=1e-12):
    # schemes.py:100-101
    return (a * (b * b + eps) + b * (a * a + eps)) / (a * a + b * b + 2.0 * eps)


def _minmod(a, b):
    # schemes.py:104-105
    return 0.5 * (np.sign(a) + np.sign(b)) * np.minimum(np.abs(a), np.abs(b))


def hllc(wl, wr, S, gamma):
    """HLLC with Davis wave speeds and Toro's star state, times |S| (schemes.py:108-137)."""
    Sm = np.sqrt(_sum3(S * S))
    n = S / Sm
    rl, ul, pl = wl[0], wl[1:4], wl[4]
    rr, ur, pr = wr[0], wr[1:4], wr[4]
    unl, unr = _sum3(ul * n), _sum3(ur * n)
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    sL, sR = np.minimum(unl - cl, unr - cr), np.maximum(unl + cl, unr + cr)
    ss = ((pr - pl + rl * unl * (sL - unl) - rr * unr * (sR - unr))
          / (rl * (sL - unl) - rr * (sR - unr)))
    El = pl / (gamma - 1.0) + 0.5 * rl * _sum3(ul * ul)
    Er = pr / (gamma - 1.0) + 0.5 * rr * _sum3(ur * ur)

    def phys(r, u, p, un, E):
        return np.concatenate([(r * un)[None], r * un * u + p * n, ((E + p) * un)[None]], 0)

    def cons(r, u, E):
        return np.concatenate([r[None], r * u, E[None]], 0)

    def star(r, u, p, un, E, s):
        f = r * (s - un) / (s - ss)
        e = E / r + (ss - un) * (ss + p / (r * (s - un)))
        return np.concatenate([f[None], f * (u + (ss - un) * n), (f * e)[None]], 0)

    Fl, Fr = phys(rl, ul, pl, unl, El), phys(rr, ur, pr, unr, Er)
    Ul, Ur = cons(rl, ul, El), cons(rr, ur, Er)
    with np.errstate(divide="ignore", invalid="ignore"):
        Fsl = Fl + sL * (star(rl, ul, pl, unl, El, sL) - Ul)
        Fsr = Fr + sR * (star(rr, ur, pr, unr, Er, sR) - Ur)
    F = np.where(sL >= 0, Fl, np.where(ss >= 0, Fsl, np.where(sR > 0, Fsr, Fr)))
    return F * Sm


class MUSCLHLLC:
    """Index-space MUSCL of primitives + positivity fallback + HLLC (schemes.py:140-152)."""

    def __init__(self, limiter="albada"):
        self.limiter = limiter
        self.lim = {"albada": _albada, "minmod": _minmod}[limiter]

    def __call__(self, blk, geom, w, gas, d, theta=None):
        S = blk.slab(geom.S[d], d, 0, 1)
        L2, L = blk.slab(w, d, -2, -1), blk.slab(w, d, -1, 0)
        R, R2 = blk.slab(w, d, 0, 1), blk.slab(w, d, 1, 2)
        wl = L + 0.5 * self.lim(L - L2, R - L)
        wr = R - 0.5 * self.lim(R - L, R2 - R)
        bad = (wl[0] <= 0) | (wl[4] <= 0) | (wr[0] <= 0) | (wr[4] <= 0)
        wl, wr = np.where(bad, L, wl), np.where(bad, R, wr)
        return hllc(wl, wr, S, gas.gamma)


class Hybrid:
    """F = (1 - s) F_central + s F_upwind, s = clip((max(th_L, th_R) - th0)/(th1 - th0))
    (schemes.py:155-166).  Both fluxes are evaluated everywhere."""

    def __init__(self, central=None, upwind=None, th0=0.6, th1=0.9):
        self.c = central or KEP(4)
        self.u = upwind or MUSCLHLLC()
        self.th0, self.th1 = th0, th1

    def __call__(self, blk, geom, w, gas, d, theta):
        th = np.maximum(blk.slab(theta, d, -1, 0), blk.slab(theta, d, 0, 1))
        s = np.clip((th - self.th0) / (self.th1 - self.th0), 0.0, 1.0)
        return (1.0 - s) * self.c(blk, geom, w, gas, d) + s * self.u(blk, geom, w, gas, d)


# ----------------------------------------------------------------------------- viscous
def green_gauss(phi, geom):
    """Cell gradients by Green-Gauss with arithmetic face averages; outermost
    ghost layer = 0 (viscous.py:6-19).  phi (nv, Ni, Nj, Nk) -> (nv, 3, Ni, Nj, Nk)."""
    nd = phi.ndim
    acc = 0.0
    for d in range(3):
        Nd = phi.shape[1 + d]
        pf = 0.5 * (phi[sl(nd, d, slice(None, -1))] + phi[sl(nd, d, slice(1, None))])
        Sf = geom.S[d][sl(4, d, slice(1, Nd))]
        flux = pf[:, None] * Sf[None]
        inner = slice(1, -1)
        acc = acc + (flux[sl(nd + 1, d, slice(1, None), inner)]
                     - flux[sl(nd + 1, d, slice(None, -1), inner)])
    g = acc / geom.vol[1:-1, 1:-1, 1:-1]
    return np.pad(g, [(0, 0), (0, 0), (1, 1), (1, 1), (1, 1)])


def ducros(gu):
    """theta = div^2 / (div^2 + |omega|^2 + 1e-30) (viscous.py:22-25)."""
    div = gu[0, 0] + gu[1, 1] + gu[2, 2]
    w2 = ((gu[2, 1] - gu[1, 2]) ** 2 + (gu[0, 2] - gu[2, 0]) ** 2 + (gu[1, 0] - gu[0, 1]) ** 2)
    return div ** 2 / (div ** 2 + w2 + 1e-30)


def viscous_flux(blk, geom, d, phi, g, mu_eff, k_eff):
    """Viscous flux (times area) through the n_d+1 interior faces of direction d
    (viscous.py:28-51): mean of the two cell gradients + compact correction along
    the cell-to-cell vector; tau = mu (grad u + grad u^T - 2/3 div I)."""
    Lf = lambda a: blk.slab(a, d, -1, 0)
    Rf = lambda a: blk.slab(a, d, 0, 1)
    gbar = 0.5 * (Lf(g) + Rf(g))
    dx = Rf(geom.xc) - Lf(geom.xc)
    dist = np.sqrt(_sum3(dx * dx))
    e = dx / dist
    ge = gbar * e[None]
    corr = (Rf(phi) - Lf(phi)) / dist - ((ge[:, 0] + ge[:, 1]) + ge[:, 2])
    gf = gbar + corr[:, None] * e[None]
    uf = 0.5 * (Lf(phi[:3]) + Rf(phi[:3]))
    muf = 0.5 * (Lf(mu_eff) + Rf(mu_eff))
    kf = 0.5 * (Lf(k_eff) + Rf(k_eff))
    S = blk.slab(geom.S[d], d, 0, 1)
    gu = gf[:3]
    div = gu[0, 0] + gu[1, 1] + gu[2, 2]
    eye = np.eye(3).reshape(3, 3, 1, 1, 1)
    tau = muf * (gu + np.swapaxes(gu, 0, 1) - (2.0 / 3.0) * eye * div)
    tauS = np.einsum('ij...,j...->i...', tau, S)
    dTdn = np.einsum('j...,j...->...', gf[3], S)
    return np.concatenate([np.zeros_like(div)[None], -tauS,
                           (-_sum3(uf * tauS) - kf * dTdn)[None]], axis=0)


# ----------------------------------------------------------------------------- SGS
def _sym(g):
    return 0.5 * (g + np.swapaxes(g, 0, 1))


def _smagorinsky(Cs=0.16):
    # sgs.py:14-18: (Cs delta)^2 sqrt(2 S:S)
    def nut(g, delta):
        S = _sym(g)
        return (Cs * delta) ** 2 * np.sqrt(2.0 * np.sum(S * S, axis=(0, 1)))
    return nut


def _wale(Cw=0.325):
    # sgs.py:21-30
    def nut(g, delta):
        S = _sym(g)
        SS = np.sum(S * S, axis=(0, 1))
        g2 = np.einsum('ik...,kj...->ij...', g, g)
        tr = g2[0, 0] + g2[1, 1] + g2[2, 2]
        Sd = _sym(g2) - np.eye(3).reshape(3, 3, 1, 1, 1) * tr / 3.0
        SdSd = np.sum(Sd * Sd, axis=(0, 1))
        return (Cw * delta) ** 2 * SdSd ** 1.5 / (SS ** 2.5 + SdSd ** 1.25 + 1e-30)
    return nut


def _vreman(Cs=0.17):
    # sgs.py:33-42
    c = 2.5 * Cs ** 2

    def nut(g, delta):
        beta = delta ** 2 * np.einsum('im...,jm...->ij...', g, g)
        B = (beta[0, 0] * beta[1, 1] - beta[0, 1] ** 2 + beta[0, 0] * beta[2, 2]
             - beta[0, 2] ** 2 + beta[1, 1] * beta[2, 2] - beta[1, 2] ** 2)
        aa = np.sum(g * g, axis=(0, 1))
        return c * np.sqrt(np.maximum(B, 0.0) / (aa + 1e-30))
    return nut


def make_sgs(name="vreman", **kw):
    """f(g, delta, rho) -> mu_t; None/'none' -> 0 (sgs.py:45-50)."""
    if name in (None, "none"):
        return lambda g, delta, rho: 0.0 * rho
    nut = {"smagorinsky": _smagorinsky, "wale": _wale, "vreman": _vreman}[name](**kw)
    return lambda g, delta, rho: rho * nut(g, delta)


# ----------------------------------------------------------------------------- wall model
class EquilibriumWallModel:
    """Spalding u_tau by safeguarded fixed-count Newton; Kader q_w (isothermal) or
    recovery-temperature adiabatic wall (wall_model.py:15-61)."""

    def __init__(self, kappa=0.41, B=5.2, n_iter=30):
        self.kappa, self.B, self.n_iter = kappa, B, n_iter

    def utau(self, h, U, rho_w, mu_w):
        k = self.kappa
        ekb = np.exp(-k * self.B)
        Reh = h * U * rho_w / mu_w

        def F(up):
            ku = k * up
            return up + ekb * (np.exp(ku) - 1.0 - ku - 0.5 * ku ** 2 - ku ** 3 / 6.0) - Reh / up

        def dF(up):
            ku = k * up
            return 1.0 + ekb * k * (np.exp(ku) - 1.0 - ku - 0.5 * ku ** 2) + Reh / up ** 2

        up = np.clip(np.sqrt(Reh), 1.0, 25.0)
        for _ in range(self.n_iter):
            new = up - F(up) / dF(up)
            up = np.minimum(np.maximum(new, 0.5 * up), 100.0)
        return U / up

    def __call__(self, gas, h, U, T_h, p_h, T_wall):
        Pr = gas.Pr
        U = np.maximum(U, 1e-8)
        if T_wall is None:
            T_w = T_h + Pr ** (1.0 / 3.0) * U ** 2 / (2.0 * gas.cp)
        else:
            T_w = T_wall + 0.0 * T_h
        rho_w = p_h / (gas.R * T_w)
        mu_w = gas.mu(T_w)
        ut = self.utau(h, U, rho_w, mu_w)
        tau_w = rho_w * ut ** 2
        if T_wall is None:
            return tau_w, 0.0 * tau_w
        yp = np.maximum(h * ut * rho_w / mu_w, 1e-6)
        Gam = np.maximum(0.01 * (Pr * yp) ** 4 / (1.0 + 5.0 * Pr ** 3 * yp), 1e-20)
        beta = (3.85 * Pr ** (1 / 3) - 1.3) ** 2 + 2.12 * np.log(Pr)
        Tp = Pr * yp * np.exp(-Gam) + (2.12 * np.log(1.0 + yp) + beta) * np.exp(-1.0 / Gam)
        q_w = rho_w * gas.cp * ut * (T_w - T_h) / Tp
        return tau_w, q_w

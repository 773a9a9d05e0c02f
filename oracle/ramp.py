"""Smooth-ramp plug-ins on the step (oracle): synthetic Fourier-mode inflow ghost fill
and spanwise-mean sponge source.  Follows reference `wmles/smooth_ramp.py:40-51,116-221`.
"""
import numpy as np

from .bc import BC
from .gas import Gas
from .mesh import sl

SPAN_L = 0.288                      # smooth_ramp.py:22


def flow_properties(Ma_inf=0.2, Re_L=667_000.0, T_inf=300.0, p_inf=101_325.0, gas=None):
    """Physical scale from free-stream Mach and Reynolds numbers (smooth_ramp.py:40-51)."""
    gas = gas or Gas()
    rho_inf = p_inf / (gas.R * T_inf)
    mu_inf = float(gas.mu(np.asarray(T_inf)))
    nu_inf = mu_inf / rho_inf
    U_inf = Ma_inf * float(gas.sound_speed(np.asarray(T_inf)))
    L = Re_L * nu_inf / U_inf
    return dict(gas=gas, Ma_inf=Ma_inf, Re_L=Re_L, T_inf=T_inf, p_inf=p_inf,
                rho_inf=rho_inf, mu_inf=mu_inf, nu_inf=nu_inf, U_inf=U_inf, L=L)


def draw_modes(length_scales, seed, n_modes, span_cells, span_l=SPAN_L):
    """Per velocity component c: default_rng(seed + c) draws kx ~ N(0, 1/Lx),
    ky ~ N(0, 1/Ly), kz from the discrete span modes with Gaussian weights, phase ~
    U(0, 2 pi) (smooth_ramp.py:131-146).  Returns a list of 3 (kx, ky, kz, phase)."""
    ls = np.asarray(length_scales, float)
    max_span_mode = max(1, span_cells // 2)
    span_modes = np.arange(-max_span_mode, max_span_mode + 1)
    span_modes = span_modes[span_modes != 0]
    kz_values = 2.0 * np.pi * span_modes / span_l
    weights = np.exp(-0.5 * (kz_values * ls[2]) ** 2)
    weights /= np.sum(weights)
    modes = []
    for comp in range(3):
        rng = np.random.default_rng(int(seed) + comp)
        kx = rng.normal(0.0, 1.0 / ls[0], n_modes)
        ky = rng.normal(0.0, 1.0 / ls[1], n_modes)
        kz = rng.choice(kz_values, size=n_modes, p=weights)
        phase = rng.uniform(0.0, 2.0 * np.pi, n_modes)
        modes.append((kx, ky, kz, phase))
    return modes


class SmoothRampInflow(BC):
    """imin ghost fill: DNS-profile mean + Fourier-mode fluctuations normalised per
    wall-normal row and shaped by a Cholesky of (uu, uv, vv) (smooth_ramp.py:116-197)."""

    def __init__(self, props, profile, length_scales=(0.1, 0.02, 0.03), seed=0, n_modes=64,
                 span_cells=16):
        self.props, self.profile = props, np.asarray(profile, float)
        self.modes = draw_modes(length_scales, seed, n_modes, span_cells)

    def fill(self, q, blk, geom, gas, axis, side):
        return self.fill_at_time(q, blk, geom, gas, axis, side, 0.0)

    def fill_at_time(self, q, blk, geom, gas, axis, side, time):
        if axis != 0 or side != 0:
            raise ValueError("SmoothRampInflow is only valid on the i-min face")
        ng = blk.ng
        q = q.copy()
        tab = self.profile
        Uref = self.props["U_inf"]
        for m in range(ng):
            gi = ng - 1 - m
            xc = geom.xc[sl(4, axis, slice(gi, gi + 1))]
            y_l, z_l = xc[1, 0] / self.props["L"], xc[2, 0] / self.props["L"]
            noise = self.unit_noise(y_l, z_l, time, blk.n[2], ng)
            interp = lambda col: np.interp(y_l, tab[:, 0], tab[:, col])
            uu = np.maximum(interp(3), 0.0) * Uref ** 2
            vv = np.maximum(interp(4), 0.0) * Uref ** 2
            ww = np.maximum(interp(5), 0.0) * Uref ** 2
            uv = interp(6) * Uref ** 2
            sqrt_uu = np.sqrt(uu)
            with np.errstate(divide="ignore", invalid="ignore"):
                l10 = np.where(sqrt_uu > 1e-14, uv / np.maximum(sqrt_uu, 1e-14), 0.0)
            l11 = np.sqrt(np.maximum(vv - l10 ** 2, 0.0))
            u = Uref * interp(1) + sqrt_uu * noise[0]
            v = Uref * interp(2) + l10 * noise[0] + l11 * noise[1]
            w = np.sqrt(ww) * noise[2]
            zero = np.zeros_like(u)
            prim = np.stack([self.props["rho_inf"] + zero, u, v, w,
                             self.props["p_inf"] + zero])[:, None, :, :]
            q[sl(q.ndim, axis, slice(gi, gi + 1))] = gas.prim_to_cons(prim)
        return q

    def unit_noise(self, y_l, z_l, time, span_cells, ng):
        """(3, Nj, Nk) unit-variance, zero-mean (over the interior span) fields;
        v is made uncorrelated with u (smooth_ramp.py:175-197)."""
        coords = (self.props["U_inf"] * time / self.props["L"], y_l, z_l)
        fields = []
        zi = slice(ng, ng + span_cells)
        for comp, (kx, ky, kz, phase) in enumerate(self.modes):
            ph = (kx[:, None, None] * coords[0] + ky[:, None, None] * coords[1][None]
                  + kz[:, None, None] * coords[2][None] + phase[:, None, None])
            field = np.sum(np.cos(ph), axis=0)
            field = field - np.mean(field[:, zi], axis=1, keepdims=True)
            if comp == 1:
                ref = fields[0]
                cov = np.mean(field[:, zi] * ref[:, zi], axis=1, keepdims=True)
                field = field - cov * ref
            var = np.mean(field[:, zi] ** 2, axis=1, keepdims=True)
            fields.append(field / np.sqrt(np.maximum(var, 1e-20)))
        return np.stack(fields)


class SmoothRampSponge:
    """Source -sigma(x) (q - <q>_z), <.>_z the volume-weighted spanwise mean,
    sigma = strength U_inf/L smoothstep5((x/L - x0)/(x1 - x0)) (smooth_ramp.py:200-221)."""

    def __init__(self, props, x_start_l, x_end_l, strength=1.0):
        self.props = props
        self.x_start_l, self.x_end_l = float(x_start_l), float(x_end_l)
        self.strength = float(strength)

    def __call__(self, blk, q, geom, gas):
        q_int = blk.interior(q)
        vol = blk.interior(geom.vol)
        mean = np.sum(q_int * vol[None], axis=3, keepdims=True) \
            / np.sum(vol, axis=2, keepdims=True)[None]
        x_l = blk.interior(geom.xc[0]) / self.props["L"]
        eta = np.clip((x_l - self.x_start_l) / (self.x_end_l - self.x_start_l), 0.0, 1.0)
        ramp = eta ** 3 * (10.0 + eta * (-15.0 + 6.0 * eta))
        sigma = self.strength * self.props["U_inf"] / self.props["L"] * ramp
        return -sigma[None] * (q_int - mean)

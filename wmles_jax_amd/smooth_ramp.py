"""Smooth-ramp plug-ins on the step (reference `wmles/smooth_ramp.py:40-51,116-221`):
the synthetic Fourier-mode inflow ghost fill and the spanwise-mean sponge source.

Mode draws use NumPy's seeded generator on the host at construction, exactly as the
reference draws them; the fill and the source run in HIP kernels.
"""
import numpy as np
import torch

from . import _lib
from .bc import BC
from .gas import Gas

RE_L = 667_000.0
MACH_INF = 0.2
RAMP_HEIGHT_L = 0.22
TOP_HEIGHT_L = 0.74
SPAN_L = 0.288


def flow_properties(Ma_inf=MACH_INF, Re_L=RE_L, T_inf=300.0, p_inf=101_325.0, gas=None):
    """Physical scale from the free-stream Mach and Reynolds numbers (smooth_ramp.py:40-51)."""
    if Ma_inf <= 0.0 or Re_L <= 0.0 or T_inf <= 0.0 or p_inf <= 0.0:
        raise ValueError("Ma_inf, Re_L, T_inf, and p_inf must all be positive")
    gas = gas or Gas()
    rho_inf = p_inf / (gas.R * T_inf)
    mu_inf = float(gas.mu(np.asarray(T_inf)))
    nu_inf = mu_inf / rho_inf
    U_inf = Ma_inf * float(gas.sound_speed(np.asarray(T_inf)))
    L = Re_L * nu_inf / U_inf
    return dict(gas=gas, Ma_inf=Ma_inf, Re_L=Re_L, T_inf=T_inf, p_inf=p_inf,
                rho_inf=rho_inf, mu_inf=mu_inf, nu_inf=nu_inf, U_inf=U_inf, L=L)


class SmoothRampInflow(BC):
    """i-min ghost fill: profile mean + unit-variance Fourier-mode fluctuations shaped by
    the Cholesky factor of (uu, uv, vv) (smooth_ramp.py:116-197)."""
    kind = _lib.BC_INFLOW

    def __init__(self, props, profile, length_scales=(0.1, 0.02, 0.03), seed=0, n_modes=64,
                 span_cells=16):
        length_scales = np.asarray(length_scales, dtype=float)
        if (length_scales.shape != (3,) or not np.all(np.isfinite(length_scales))
                or np.any(length_scales <= 0.0)):
            raise ValueError("inflow length_scales must contain three positive finite values")
        if seed < 0:
            raise ValueError("inflow seed must be non-negative")
        if n_modes < 4 or span_cells < 3:
            raise ValueError("inflow synthesis needs at least four modes and three spanwise cells")
        self.props, self.profile = props, np.asarray(profile, dtype=float)
        self.length_scales = tuple(length_scales)
        self.seed, self.n_modes = int(seed), int(n_modes)
        max_span_mode = max(1, span_cells // 2)
        span_modes = np.arange(-max_span_mode, max_span_mode + 1)
        span_modes = span_modes[span_modes != 0]
        kz_values = 2.0 * np.pi * span_modes / SPAN_L
        weights = np.exp(-0.5 * (kz_values * length_scales[2]) ** 2)
        weights /= np.sum(weights)
        modes = np.empty((3, 4, self.n_modes))
        for comp in range(3):
            rng = np.random.default_rng(self.seed + comp)
            modes[comp, 0] = rng.normal(0.0, 1.0 / length_scales[0], self.n_modes)
            modes[comp, 1] = rng.normal(0.0, 1.0 / length_scales[1], self.n_modes)
            modes[comp, 2] = rng.choice(kz_values, size=self.n_modes, p=weights)
            modes[comp, 3] = rng.uniform(0.0, 2.0 * np.pi, self.n_modes)
        self.modes = modes
        self._dev_arrays = None

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.BC_INFLOW
        if ctx is not None and hasattr(ctx, "inflows"):
            if self not in ctx.inflows:
                ctx.inflows.append(self)
            fd.inflow_id = ctx.inflows.index(self)

    def desc(self, dev):
        if self._dev_arrays is None:
            tab = torch.as_tensor(np.ascontiguousarray(self.profile), dtype=torch.float64, device=dev)
            mod = torch.as_tensor(np.ascontiguousarray(self.modes), dtype=torch.float64, device=dev)
            self._dev_arrays = (tab, mod)
        tab, mod = self._dev_arrays
        d = _lib.InflowDesc()
        d.n_tab = int(self.profile.shape[0])
        d.table = tab.data_ptr()
        d.n_modes = self.n_modes
        d.modes = mod.data_ptr()
        d.U_inf, d.L = float(self.props["U_inf"]), float(self.props["L"])
        d.rho_inf, d.p_inf = float(self.props["rho_inf"]), float(self.props["p_inf"])
        return d


class SmoothRampSponge:
    """Source -sigma(x) (q - <q>_z), sigma = strength U_inf/L smoothstep5 over
    [x_start_l, x_end_l] (smooth_ramp.py:200-221)."""

    def __init__(self, props, x_start_l, x_end_l, strength=1.0):
        if not np.isfinite(strength) or strength < 0.0:
            raise ValueError("sponge strength must be finite and non-negative")
        if not np.isfinite(x_start_l + x_end_l) or x_start_l >= x_end_l:
            raise ValueError("sponge start must be finite and upstream of its end")
        self.props = props
        self.x_start_l, self.x_end_l = float(x_start_l), float(x_end_l)
        self.strength = float(strength)

    def __call__(self, blk, q, geom, gas):
        raise NotImplementedError("the sponge source is evaluated inside the HIP step")

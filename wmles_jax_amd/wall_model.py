"""Equilibrium wall model (reference `wmles/wall_model.py:15-61`): Spalding's law by a
safeguarded fixed-count Newton solve for u_tau, Kader's temperature law for q_w, or a
recovery-temperature adiabatic wall.  Evaluated by the HIP wall-flux kernel."""
import ctypes as C

import numpy as np
import torch

from . import _lib


class EquilibriumWallModel:
    def __init__(self, kappa=0.41, B=5.2, n_iter=30):
        self.kappa, self.B, self.n_iter = kappa, B, n_iter

    def host_constants(self, gas):
        """exp(-kappa B), Pr^(1/3), Pr^3, Kader beta, evaluated as the reference does
        on the host (wall_model.py:21, 35, 45-48)."""
        Pr = gas.Pr
        ekb = float(np.exp(-self.kappa * self.B))
        pr13 = Pr ** (1.0 / 3.0)
        pr3 = Pr ** 3
        beta = float((3.85 * Pr ** (1 / 3) - 1.3) ** 2 + 2.12 * np.log(Pr))
        return ekb, pr13, pr3, beta

    def _eval(self, mode, a, b, c, d, T_wall=None, gas=None):
        from .backend import asarray
        from .gas import Gas
        from .solver import numerics_desc
        gas = gas or Gas()
        nd = numerics_desc(gas, None, None, 0.8, 2, (), ())
        nd.kappa, nd.B, nd.n_iter = self.kappa, self.B, self.n_iter
        nd.wm_ekb, nd.wm_pr13, nd.wm_pr3, nd.wm_beta = self.host_constants(gas)
        ts = [asarray(x).reshape(-1) for x in (a, b, c, d)]
        n = max(t.numel() for t in ts)
        ts = [t.expand(n).contiguous() if t.numel() == 1 else t for t in ts]
        o0 = torch.empty(n, dtype=torch.float64, device=ts[0].device)
        o1 = torch.empty_like(o0)
        tw = -1.0 if T_wall is None else float(T_wall)
        _lib.check(_lib.lib.wmx_wall_model_eval(C.byref(nd), mode, n, *[t.data_ptr() for t in ts],
                                                tw, o0.data_ptr(), o1.data_ptr(), _lib.stream_ptr()))
        shape = tuple(np.broadcast(*[np.empty(np.shape(x)) for x in (a, b, c, d)]).shape)
        return o0.reshape(shape), o1.reshape(shape)

    def utau(self, h, U, rho_w, mu_w):
        """Spalding u_tau by the safeguarded fixed-count Newton (wall_model.py:19-40)."""
        return self._eval(0, h, U, rho_w, mu_w)[0]

    def __call__(self, gas, h, U, T_h, p_h, T_wall):
        """(tau_w, q_w) (wall_model.py:42-61)."""
        return self._eval(1, h, U, T_h, p_h, T_wall, gas)

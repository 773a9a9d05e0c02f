"""Calorically perfect gas with Sutherland viscosity (reference `wmles/gas.py:10-41`).

The step itself evaluates these relations inside the HIP kernels
(`csrc/physics.h`); the methods here serve setup (`Solver.init_state`) and
post-processing and accept NumPy arrays or torch tensors alike.
"""
from dataclasses import dataclass

import numpy as np


def _cat(parts):
    if hasattr(parts[0], "device") and not isinstance(parts[0], np.ndarray):
        import torch
        return torch.cat(parts, dim=0)
    return np.concatenate(parts, axis=0)


def _sqrt(a):
    if isinstance(a, (float, int, np.ndarray, np.floating)):
        return np.sqrt(a)
    return a.sqrt()


@dataclass(frozen=True)
class Gas:
    gamma: float = 1.4
    R: float = 287.05
    Pr: float = 0.72
    Pr_t: float = 0.9
    mu_ref: float = 1.716e-5
    T_ref: float = 273.15
    S_suth: float = 110.4

    @property
    def cp(self):
        return self.gamma * self.R / (self.gamma - 1.0)

    def mu(self, T):
        return (self.mu_ref * (T / self.T_ref) ** 1.5
                * (self.T_ref + self.S_suth) / (T + self.S_suth))

    def cons_to_prim(self, q):
        rho = q[0]
        vel = q[1:4] / rho
        k = (vel[0] * vel[0] + vel[1] * vel[1]) + vel[2] * vel[2]
        p = (self.gamma - 1.0) * (q[4] - 0.5 * rho * k)
        T = p / (rho * self.R)
        return _cat([rho[None], vel, p[None]]), T

    def prim_to_cons(self, w):
        rho, vel, p = w[0], w[1:4], w[4]
        k = (vel[0] * vel[0] + vel[1] * vel[1]) + vel[2] * vel[2]
        E = p / (self.gamma - 1.0) + 0.5 * rho * k
        return _cat([rho[None], rho * vel, E[None]])

    def sound_speed(self, T):
        return _sqrt(self.gamma * self.R * T)

"""Perfect gas + Sutherland viscosity (oracle).  Follows reference `wmles/gas.py:10-41`.

Conservative q = [rho, rho u, rho v, rho w, rho E]; primitive w = [rho, u, v, w, p].
Arithmetic is written in the same association order as the reference so the oracle
reproduces its float64 values.
"""
from dataclasses import dataclass

import numpy as np


def _sum3(a):
    """Sum over the leading length-3 axis in index order: (a0 + a1) + a2."""
    return (a[0] + a[1]) + a[2]


@dataclass(frozen=True)
class Gas:
    """Reference `gas.py:10-18` defaults."""
    gamma: float = 1.4
    R: float = 287.05
    Pr: float = 0.72
    Pr_t: float = 0.9
    mu_ref: float = 1.716e-5
    T_ref: float = 273.15
    S_suth: float = 110.4

    @property
    def cp(self):
        # gas.py:20-22
        return self.gamma * self.R / (self.gamma - 1.0)

    def mu(self, T):
        # gas.py:24-26: mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S)
        return (self.mu_ref * (T / self.T_ref) ** 1.5
                * (self.T_ref + self.S_suth) / (T + self.S_suth))

    def cons_to_prim(self, q):
        # gas.py:28-33
        rho = q[0]
        vel = q[1:4] / rho
        p = (self.gamma - 1.0) * (q[4] - 0.5 * rho * _sum3(vel * vel))
        T = p / (rho * self.R)
        return np.concatenate([rho[None], vel, p[None]], axis=0), T

    def prim_to_cons(self, w):
        # gas.py:35-38
        rho, vel, p = w[0], w[1:4], w[4]
        E = p / (self.gamma - 1.0) + 0.5 * rho * _sum3(vel * vel)
        return np.concatenate([rho[None], rho * vel, E[None]], axis=0)

    def sound_speed(self, T):
        # gas.py:40-41
        return np.sqrt(self.gamma * self.R * T)

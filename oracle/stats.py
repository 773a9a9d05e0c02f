"""CPU restatement of the reference's channel statistics -- TEST INFRASTRUCTURE ONLY.

`PlaneStats` follows `wmles/channel5200.py:121-199`: per accumulate, the (x, z)-plane
means (axes 0 and 2 of the interior) of rho, u, v, w, uu, vv, ww, uv, uw, vw, T and p,
plus the wall-mean tau_w; result() divides by the sample count and subtracts the mean
products from the second moments; wall_units() scales the lower half by u_tau.

Pinned by closed forms on the CPU (tests/test_oracle_stats.py): a state whose velocity
is f(y) + A sin(2 pi i / n_x) on whole periods has plane mean f(y) and uu' = A^2 / 2,
uv' = 0, T = p / (rho R); a pure-Python triple loop on a small random state; the
Reynolds-stress subtraction against two samples whose combined variance is known.
"""
import numpy as np

KEYS = ("rho", "u", "v", "w", "uu", "vv", "ww", "uv", "uw", "vw", "T", "p")


def plane_means(q_interior, gas):
    """(12, Ny) plane means of one interior state (5, Nx, Ny, Nz) (channel5200.py:139-157)."""
    w, T = gas.cons_to_prim(q_interior)
    rho, u, v, wc, p = w[0], w[1], w[2], w[3], w[4]
    ax = (0, 2)
    vals = (rho, u, v, wc, u * u, v * v, wc * wc, u * v, u * wc, v * wc, T, p)
    return np.stack([x.mean(axis=ax) for x in vals])


class PlaneStats:
    """channel5200.py:121-199 on interior states given directly (block 0 only)."""

    def __init__(self, y):
        self.y = np.asarray(y, dtype=np.float64)
        self.n = 0
        self.tau_w_sum = 0.0
        self.sum = {k: np.zeros_like(self.y) for k in KEYS}

    def accumulate(self, q_interior, gas, tau_w_mean):
        m = plane_means(q_interior, gas)
        for i, k in enumerate(KEYS):
            self.sum[k] += m[i]
        self.tau_w_sum += tau_w_mean
        self.n += 1

    def result(self):
        avg = {k: v / self.n for k, v in self.sum.items()}
        avg["uu"] -= avg["u"] ** 2
        avg["vv"] -= avg["v"] ** 2
        avg["ww"] -= avg["w"] ** 2
        avg["uv"] -= avg["u"] * avg["v"]
        avg["uw"] -= avg["u"] * avg["w"]
        avg["vw"] -= avg["v"] * avg["w"]
        return self.y, avg

    @property
    def mean_tau_w(self):
        return self.tau_w_sum / self.n

    def wall_units(self, tau_w, rho_w, mu_w, h):
        y, avg = self.result()
        u_tau = np.sqrt(tau_w / rho_w)
        nu_w = mu_w / rho_w
        mask = y < h
        y, yp = y[mask], y[mask] * u_tau / nu_w
        u_plus = avg["u"][mask] / u_tau
        p = avg["p"][mask]
        p_wall = p[0] - y[0] * (p[1] - p[0]) / (y[1] - y[0])
        out = {"y/delta": y / h, "y+": yp, "U+": u_plus, "dU/dy+": np.gradient(u_plus, yp),
               "W+": avg["w"][mask] / u_tau, "P+": (p - p_wall) / (rho_w * u_tau ** 2)}
        for k in ("uu", "vv", "ww", "uv", "uw", "vw"):
            out[k + "+"] = avg[k][mask] / u_tau ** 2
        out["k+"] = 0.5 * (out["uu+"] + out["vv+"] + out["ww+"])
        return u_tau, out

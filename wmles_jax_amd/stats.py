"""Plane- and time-averaged channel statistics on the device.

`PlaneStats(mesh, solver)` mirrors the reference's statistics object
(`wmles/channel5200.py:121-199`): `accumulate(Q)` adds the (x, z)-plane means of
rho, u, v, w, the velocity products, T and p of block 0 (the HIP kernel
`wmx_plane_stats`, FP64, one launch pair per call; nothing leaves the device) and the
wall-mean tau_w; `result()` returns (y, profiles) with the Reynolds stresses formed
from the time means; `wall_units(...)` the inner-scaled lower-half profiles.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

KEYS = ("rho", "u", "v", "w", "uu", "vv", "ww", "uv", "uw", "vw", "T", "p")


class PlaneStats:
    """Streamwise/spanwise-plane- and time-averaged profiles of block 0 (an unskewed
    channel block: Ny distinct y planes shared by every (x, z))."""

    def __init__(self, mesh, solver):
        self.mesh, self.solver = mesh, solver
        blk, geom = mesh.blocks[0], solver.geoms[0]
        self.y = np.asarray(torch.as_tensor(blk.interior(geom.xc[1])).cpu())[0, :, 0]
        self.n = 0
        self.tau_w_sum = 0.0
        self._acc = torch.zeros((len(KEYS), blk.n[1]), dtype=torch.float64, device=solver._dev)

    def accumulate(self, Q):
        Q = self.solver._as_state(Q)
        _lib.check(_lib.lib.wmx_plane_stats(self.solver._plan, _lib.ptr_array(Q), 0,
                                            C.c_void_p(self._acc.data_ptr()), _lib.stream_ptr()))
        wall = self.solver.wall_stats(Q)
        self.tau_w_sum += float(np.mean([np.mean(v[0]) for v in wall.values()]))
        self.n += 1

    @property
    def sum(self):
        """Accumulated plane means, {name: (Ny,) numpy array}."""
        a = self._acc.cpu().numpy()
        return {k: a[i].copy() for i, k in enumerate(KEYS)}

    def result(self):
        if self.n <= 0:
            raise AssertionError("call accumulate() at least once before result()")
        avg = {k: v / self.n for k, v in self.sum.items()}
        for a, b in (("u", "u"), ("v", "v"), ("w", "w"), ("u", "v"), ("u", "w"), ("v", "w")):
            avg[a + b] = avg[a + b] - avg[a] * avg[b]
        return self.y, avg

    @property
    def mean_tau_w(self):
        """Mean wall shear stress over the accumulated samples."""
        if self.n <= 0:
            raise AssertionError("call accumulate() at least once before mean_tau_w")
        return self.tau_w_sum / self.n

    def wall_units(self, tau_w, rho_w, mu_w, h):
        """Inner-scaled profiles of the lower half (y < h): returns (u_tau, dict)."""
        y, avg = self.result()
        u_tau = float(np.sqrt(tau_w / rho_w))
        nu_w = mu_w / rho_w
        lower = y < h
        y = y[lower]
        yp = y * u_tau / nu_w
        u_plus = avg["u"][lower] / u_tau
        p = avg["p"][lower]
        p_wall = p[0] - y[0] * (p[1] - p[0]) / (y[1] - y[0])     # linear extrapolation to y = 0
        ut2 = u_tau * u_tau
        out = {"y/delta": y / h, "y+": yp, "U+": u_plus, "dU/dy+": np.gradient(u_plus, yp),
               "W+": avg["w"][lower] / u_tau, "P+": (p - p_wall) / (rho_w * ut2)}
        for k in ("uu", "vv", "ww", "uv", "uw", "vw"):
            out[k + "+"] = avg[k][lower] / ut2
        out["k+"] = 0.5 * (out["uu+"] + out["vv+"] + out["ww+"])
        return u_tau, out

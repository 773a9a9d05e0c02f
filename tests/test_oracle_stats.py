"""Pins the oracle's channel statistics (oracle/stats.py) by closed forms and brute force."""
import math

import numpy as np

import oracle
from oracle import stats as OS


def _state(u, v, w, rho, p, gas):
    prim = np.stack([rho, u, v, w, p])
    return gas.prim_to_cons(prim)


def test_plane_means_closed_form():
    """u = f(y) + A sin(2 pi i / nx), v = B cos(2 pi k / nz), w = 0 on whole periods:
    plane means f(y), 0, 0; <u'u'> = A^2/2, <v'v'> = B^2/2, <u'v'> = 0; T = p/(rho R)."""
    gas = oracle.Gas()
    nx, ny, nz = 12, 5, 8
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[None, None, :]
    f = 10.0 + 3.0 * j + 0.0 * i + 0.0 * k
    A, B = 2.5, 1.5
    u = f + A * np.sin(2 * math.pi * i / nx) + 0.0 * k
    v = B * np.cos(2 * math.pi * k / nz) + 0.0 * i + 0.0 * j
    w = 0.0 * u
    rho = 1.2 + 0.0 * u
    p = 1e5 + 0.0 * u
    q = _state(u, v, w, rho, p, gas)
    st = OS.PlaneStats(np.arange(ny, dtype=float))
    st.accumulate(q, gas, 2.0)
    y, avg = st.result()
    np.testing.assert_allclose(avg["u"], f[0, :, 0], rtol=1e-14)
    np.testing.assert_allclose(avg["v"], 0.0, atol=1e-13)
    np.testing.assert_allclose(avg["uu"], A * A / 2, rtol=1e-12)
    np.testing.assert_allclose(avg["vv"], B * B / 2, rtol=1e-12)
    np.testing.assert_allclose(avg["uv"], 0.0, atol=1e-12)
    np.testing.assert_allclose(avg["ww"], 0.0, atol=1e-20)
    np.testing.assert_allclose(avg["T"], 1e5 / (1.2 * gas.R), rtol=1e-13)
    np.testing.assert_allclose(avg["p"], 1e5, rtol=1e-13)
    assert st.mean_tau_w == 2.0


def test_plane_means_brute_force_and_two_sample_variance():
    """Triple-loop means of a small random state; two samples u and u + d (uniform d)
    give a time-mean Reynolds stress of d^2/4 plus the spatial part."""
    gas = oracle.Gas()
    rng = np.random.default_rng(5)
    nx, ny, nz = 3, 4, 5
    rho = 1.0 + 0.1 * rng.random((nx, ny, nz))
    u, v, w = (rng.standard_normal((nx, ny, nz)) for _ in range(3))
    p = 1e5 * (1.0 + 0.01 * rng.random((nx, ny, nz)))
    q = _state(u, v, w, rho, p, gas)
    m = OS.plane_means(q, gas)
    for jj in range(ny):
        acc = np.zeros(12)
        for ii in range(nx):
            for kk in range(nz):
                r, uu, vv, ww, pp = rho[ii, jj, kk], u[ii, jj, kk], v[ii, jj, kk], w[ii, jj, kk], p[ii, jj, kk]
                acc += [r, uu, vv, ww, uu * uu, vv * vv, ww * ww, uu * vv, uu * ww, vv * ww, pp / (r * gas.R), pp]
        np.testing.assert_allclose(m[:, jj], acc / (nx * nz), rtol=1e-9)
    d = 0.7
    st = OS.PlaneStats(np.arange(ny, dtype=float))
    st.accumulate(q, gas, 1.0)
    st.accumulate(_state(u + d, v, w, rho, p, gas), gas, 3.0)
    _, avg = st.result()
    spatial = (u * u).mean(axis=(0, 2)) - u.mean(axis=(0, 2)) ** 2
    np.testing.assert_allclose(avg["uu"], spatial + d * d / 4, rtol=1e-9)
    assert st.mean_tau_w == 2.0


def test_wall_units_log_law_profile():
    """A log-law mean profile U = u_tau (ln(y u_tau / nu) / kappa + B) comes back as U+."""
    gas = oracle.Gas()
    ny = 16
    y = np.linspace(0.05, 1.95, ny)
    rho_w, mu_w, tau_w, h = 1.2, 1.8e-5, 4.0, 1.0
    ut = math.sqrt(tau_w / rho_w)
    U = ut * (np.log(y * ut * rho_w / mu_w) / 0.41 + 5.2)
    u = np.broadcast_to(U[None, :, None], (4, ny, 3)).copy()
    z = 0.0 * u
    q = _state(u, z, z, 1.2 + z, 1e5 + z, gas)
    st = OS.PlaneStats(y)
    st.accumulate(q, gas, tau_w)
    u_tau, out = st.wall_units(st.mean_tau_w, rho_w, mu_w, h)
    assert abs(u_tau - ut) < 1e-15
    np.testing.assert_allclose(out["U+"], np.log(out["y+"]) / 0.41 + 5.2, rtol=1e-12)
    np.testing.assert_allclose(out["P+"], 0.0, atol=1e-9)
    assert out["y/delta"].max() < 1.0

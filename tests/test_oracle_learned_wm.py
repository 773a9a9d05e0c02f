"""Pins of the learned wall model's CPU oracle (training data, network gradient, Adam)."""
import numpy as np
import pytest
import torch

from oracle import learned_wm as L
from oracle.numerics import EquilibriumWallModel


def _splitmix64_int(state):
    """Vigna's splitmix64 on Python integers: (next state, output)."""
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return state, z ^ (z >> 31)


def test_splitmix64_published_sequence():
    # splitmix64 seeded with 0: the first outputs of the published generator
    s, out = 0, []
    for _ in range(3):
        s, z = _splitmix64_int(s)
        out.append(z)
    assert out[0] == 0xE220A8397B1DCDAF
    assert out[1] == 0x6E789E6AA1B965F4
    assert out[2] == 0x06C45D188009454F


def test_uhash_is_splitmix64_of_counter():
    # uhash(seed, i, s) = splitmix64 output at state seed ^ (g (4 i + s + 1)), top 53 bits
    g, m = 0x9E3779B97F4A7C15, (1 << 64) - 1
    for seed in (0, 7, 123456789):
        for i in (0, 1, 5, 1000003):
            for s in (0, 1, 2):
                st = seed ^ ((g * (4 * i + s + 1)) & m)
                _, z = _splitmix64_int(st)
                assert L.uhash(seed, np.array([i]), s)[0] == (z >> 11) / 2.0 ** 53


def test_dataset_targets_satisfy_the_laws():
    lo, hi = (np.log(1e2), 0.05, 0.5), (np.log(1e6), 4.0, 1.5)
    xc = [0.5 * (a + b) for a, b in zip(lo, hi)]
    xs = [0.5 * (b - a) for a, b in zip(lo, hi)]
    X, Y, raw = L.dataset(4096, 3, lo, hi, xc, xs, (0.0, 0.0), (1.0, 1.0))
    f, up, Tp = raw[:, :3], raw[:, 3], raw[:, 4]
    assert np.all(f >= np.array(lo)) and np.all(f < np.array(hi))
    # uniform coverage of each feature
    for k in range(3):
        u = (f[:, k] - lo[k]) / (hi[k] - lo[k])
        assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    # Spalding's law holds at the Newton solution: u+ + E(u+) = Re_h / u+
    Reh = np.exp(f[:, 0])
    k, B = 0.41, 5.2
    ku = k * up
    E = np.exp(-k * B) * (np.expm1(ku) - ku - ku ** 2 / 2 - ku ** 3 / 6)
    # (relative to the largest term, Re_h / u+: the residual is at roundoff)
    assert np.max(np.abs(up + E - Reh / up) / (Reh / up)) < 1e-14
    # log law far out, viscous sublayer limit near the wall (u+ ~ y+ = Re_h / u+)
    yp = Reh / up
    far = yp > 2000
    assert np.max(np.abs(up[far] - (np.log(yp[far]) / k + B))) < 0.05
    # Kader at y+: independent evaluation of the closed form with Pr = 0.72
    Pr = 0.72
    Gam = 0.01 * (Pr * yp) ** 4 / (1 + 5 * Pr ** 3 * yp)
    beta = (3.85 * Pr ** (1 / 3) - 1.3) ** 2 + 2.12 * np.log(Pr)
    T_ref = Pr * yp * np.exp(-Gam) + (2.12 * np.log1p(yp) + beta) * np.exp(-1 / Gam)
    np.testing.assert_allclose(Tp, T_ref, rtol=1e-12)
    np.testing.assert_allclose(Y[:, 0], np.log(up).astype(np.float32), rtol=1e-6)


def test_uplus_matches_bisection():
    Reh = np.geomspace(1e2, 1e6, 64)
    up = L.uplus_newton(Reh)
    k, B = 0.41, 5.2

    def F(u):
        ku = k * u
        return u + np.exp(-k * B) * (np.exp(ku) - 1 - ku - ku ** 2 / 2 - ku ** 3 / 6) - Reh / u

    a, b = np.full_like(Reh, 1e-3), np.full_like(Reh, 100.0)
    for _ in range(200):
        c = 0.5 * (a + b)
        neg = F(c) < 0
        a, b = np.where(neg, c, a), np.where(neg, b, c)
    np.testing.assert_allclose(up, 0.5 * (a + b), rtol=1e-12)


def _problem(seed=0, n=300):
    rng = np.random.default_rng(seed)
    P = L.init_params(seed).astype(np.float64) + 0.05 * rng.standard_normal(L.NPARAM)
    X = rng.uniform(-1, 1, (n, 3))
    Y = rng.uniform(-1, 1, (n, 2))
    return P, X, Y


def test_gradient_matches_finite_differences():
    P, X, Y = _problem(1, 97)
    _, g = L.loss_grad(P, X, Y)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(np.arange(s, e), min(6, e - s), replace=False)
                          for s, e in ((L.O_W1, L.O_B1), (L.O_B1, L.O_W2), (L.O_W2, L.O_B2),
                                       (L.O_B2, L.O_W3), (L.O_W3, L.O_B3), (L.O_B3, L.NPARAM))])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (L.loss_grad(Pp, X, Y)[0] - L.loss_grad(Pm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_gradient_matches_torch_autograd():
    P, X, Y = _problem(2, 200)
    loss, g = L.loss_grad(P, X, Y)
    Pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    W1 = Pt[L.O_W1:L.O_B1].view(64, 3)
    b1 = Pt[L.O_B1:L.O_W2]
    W2 = Pt[L.O_W2:L.O_B2].view(64, 64)
    b2 = Pt[L.O_B2:L.O_W3]
    W3 = Pt[L.O_W3:L.O_B3].view(2, 64)
    b3 = Pt[L.O_B3:]
    Xt, Yt = torch.tensor(X), torch.tensor(Y)
    act = lambda a: a * torch.rsqrt(1.0 + a * a)
    out = act(act(Xt @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    lt = torch.mean((out - Yt) ** 2)
    lt.backward()
    assert abs(float(lt.detach()) - loss) < 1e-14
    np.testing.assert_allclose(g, Pt.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_activation_and_derivative():
    a = np.linspace(-6, 6, 241)
    h = L.act(a)
    np.testing.assert_allclose(h, a / np.sqrt(1 + a * a), rtol=1e-15)
    assert np.all(np.abs(h) < 1) and np.all(np.diff(h) > 0)           # bounded, increasing
    eps = 1e-5                                    # central difference: O(eps^2) + roundoff/eps
    fd = (L.act(a + eps) - L.act(a - eps)) / (2 * eps)
    np.testing.assert_allclose(L.act_prime(h), fd, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(L.act_prime(h), (1 + a * a) ** -1.5, rtol=1e-12)


def test_adam_first_step_is_signed_rate():
    rng = np.random.default_rng(3)
    P = rng.standard_normal(50)
    g = rng.standard_normal(50)
    g[np.abs(g) < 0.1] = 0.5
    P1, m, v = L.adam(P, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    np.testing.assert_allclose(P1, P - 1e-3 * np.sign(g), rtol=0, atol=1e-9)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_newton_oracle_reused_by_dataset():
    # u+ of the data generator is the reference Newton's fixed-iteration result
    Reh = np.array([150.0, 3e3, 7e5])
    up = L.uplus_newton(Reh)
    wm = EquilibriumWallModel()
    np.testing.assert_allclose(wm.utau(1.0, 1.0, Reh, 1.0), 1.0 / up, rtol=1e-15)

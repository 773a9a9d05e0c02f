"""Learned wall model on the device: MFMA lane layout, training data and gradients
against the CPU oracle, the trained network's accuracy, the polished model against the
reference Newton (the judge of its error), what the network gains, and the model
inside the solver's wall kernel."""
import math
import time

import numpy as np
import pytest

import cases
import oracle
from oracle import learned_wm as OL

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402
from wmles_jax_amd import _lib  # noqa: E402

DEV = torch.device("cuda")


@pytest.fixture(scope="module")
def trained():
    m = wmx.LearnedWallModel()
    t0 = time.perf_counter()
    hist = m.train(n_samples=1 << 19, batch=8192, steps=20000, lr=2e-3, seed=0)
    print(f"[learned-wm] trained 20000 steps in {time.perf_counter() - t0:.2f} s, "
          f"final mse {hist['final_mse']:.3e}", flush=True)
    return m


def _bf16(a):
    return torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()


def test_mfma_bf16_lane_layout():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 16)).astype(np.float32)
    B = rng.standard_normal((16, 16)).astype(np.float32)
    D = torch.empty((16, 16), dtype=torch.float32, device=DEV)
    Ad, Bd = torch.tensor(A, device=DEV), torch.tensor(B, device=DEV)
    _lib.check(_lib.lib.wmx_mfma_bf16_probe(Ad.data_ptr(), Bd.data_ptr(), D.data_ptr(), _lib.stream_ptr()))
    ref = _bf16(A) @ _bf16(B)
    scale = np.abs(_bf16(A)) @ np.abs(_bf16(B))
    assert np.all(np.abs(D.cpu().numpy() - ref) <= 1e-6 * scale + 1e-30)


def test_dataset_matches_oracle():
    m = wmx.LearnedWallModel()
    m.y_center, m.y_scale = (2.9, 2.6), (0.6, 0.8)
    X, Y, R = m.dataset(wmx.Gas(), 6000, seed=11, raw=True)
    Xo, Yo, Ro = OL.dataset(6000, 11, m.lo, m.hi, m.x_center, m.x_scale, m.y_center, m.y_scale)
    R = R.cpu().numpy()
    np.testing.assert_allclose(R[:, :3], Ro[:, :3], rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(R[:, 3], Ro[:, 3], rtol=1e-13)
    np.testing.assert_allclose(R[:, 4], Ro[:, 4], rtol=1e-12)
    np.testing.assert_allclose(X.cpu().numpy(), Xo, rtol=0, atol=3e-7)
    np.testing.assert_allclose(Y.cpu().numpy(), Yo, rtol=0, atol=3e-7)


def _problem(seed, n):
    rng = np.random.default_rng(seed)
    P = (OL.init_params(seed) + 0.05 * rng.standard_normal(OL.NPARAM)).astype(np.float32)
    X = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    Y = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    return P, X, Y


@pytest.mark.parametrize("offset,batch", [(0, 1000), (37, 130), (0, 64), (5, 1), (100, 8192)])
def test_gradient_matches_oracle(offset, batch):
    """FP32 forward + backward kernels against the float64 oracle gradient (itself
    pinned by finite differences and torch.autograd)."""
    P, X, Y = _problem(1, 8400)
    m = wmx.LearnedWallModel()
    m.set_params(P)
    g, loss = m.grad(torch.tensor(X, device=DEV), torch.tensor(Y, device=DEV), offset, batch)
    lo, go = OL.loss_grad(P.astype(np.float64), X[offset:offset + batch], Y[offset:offset + batch])
    g = g.cpu().numpy().astype(np.float64)
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"[learned-wm] grad offset {offset} batch {batch}: rel err {err:.2e}", flush=True)
    assert err < 2e-5
    assert abs(float(loss) - lo) <= 1e-5 * lo


def test_forward32_matches_oracle():
    P, X, _ = _problem(2, 5000)
    m = wmx.LearnedWallModel()
    m.set_params(P)
    Y = m.forward32(torch.tensor(X, device=DEV)).cpu().numpy()
    Yo, _ = OL.forward(P.astype(np.float64), X)
    assert np.abs(Y - Yo).max() < 2e-5


def test_adam_update_matches_oracle():
    P, X, Y = _problem(3, 1024)
    m = wmx.LearnedWallModel()
    m.set_params(P)
    Xd, Yd = torch.tensor(X, device=DEV), torch.tensor(Y, device=DEV)
    mm = torch.zeros(OL.NPARAM, dtype=torch.float32, device=DEV)
    vv = torch.zeros_like(mm)
    Po, mo, vo = P.astype(np.float64), np.zeros(OL.NPARAM), np.zeros(OL.NPARAM)
    for step, lr in ((1, 1e-3), (2, 5e-4), (3, 5e-4)):
        g, _ = m.grad(Xd, Yd, 0, 1024)                 # the gradient the step will use
        m.adam_step(Xd, Yd, 0, 1024, mm, vv, step, lr)
        Po, mo, vo = OL.adam(Po, g.cpu().numpy().astype(np.float64), mo, vo, step, lr)
        np.testing.assert_allclose(m.params.cpu().numpy(), Po, rtol=0, atol=2e-6)


def test_training_reduces_loss_and_learns_the_law(trained):
    m = trained
    h = m.history
    # raw network (mode 1) against the Newton oracle over the trained range
    gas = wmx.Gas()
    rng = np.random.default_rng(7)
    n = 20000
    Reh = np.exp(rng.uniform(math.log(1e2), math.log(1e6), n))
    U, T_w, p = 10.0, 300.0, 1e5
    rho_w = p / (gas.R * T_w)
    mu_w = float(oracle.Gas().mu(T_w))
    hgt = Reh * mu_w / (U * rho_w)
    Th = 300.0 * rng.uniform(0.8, 1.2, n)
    tau, qw, up = m.eval(gas, torch.tensor(hgt, device=DEV), U, torch.tensor(Th, device=DEV), p,
                         T_wall=T_w, mode=1)
    up_ref = OL.uplus_newton(Reh)
    err = np.abs(up.cpu().numpy() / up_ref - 1.0)
    to, qo = oracle.EquilibriumWallModel()(oracle.Gas(), hgt, U, Th, p, T_w)
    qerr = np.abs(qw.cpu().numpy() - qo) / np.abs(qo).max()
    print(f"[learned-wm] raw network u+ rel err max {err.max():.2e} mean {err.mean():.2e}; "
          f"q_w rel err max {qerr.max():.2e}", flush=True)
    assert h["final_mse"] < 5e-5, h["final_mse"]
    assert err.max() < 2e-2 and err.mean() < 5e-3
    assert qerr.max() < 5e-2


def _wall_points(n, seed, out_of_range=True):
    rng = np.random.default_rng(seed)
    gas = oracle.Gas()
    lo, hi = (50.0, 2e6) if out_of_range else (1.2e2, 8e5)
    Reh = np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    U = rng.uniform(5.0, 80.0, n)
    Th = rng.uniform(250.0, 350.0, n)
    p = rng.uniform(5e4, 2e5, n)
    T_w = 300.0
    rho_w = p / (gas.R * T_w)
    h = Reh * gas.mu(T_w) / (U * rho_w)
    return h, U, Th, p, T_w


@pytest.mark.parametrize("T_wall", [300.0, None])
def test_polished_model_equals_reference_newton(trained, T_wall):
    """Mode 0 (network start + Newton polish + residual check + fallback) reproduces
    the reference model's (tau_w, q_w) to roundoff, in and out of the trained range."""
    h, U, Th, p, T_w = _wall_points(40000, 3)
    Tw = T_wall
    m = trained
    m.fallbacks.zero_()
    tau, qw, _ = m.eval(wmx.Gas(), *(torch.tensor(a, device=DEV) for a in (h, U, Th, p)), T_wall=Tw)
    to, qo = oracle.EquilibriumWallModel()(oracle.Gas(), h, U, Th, p, Tw)
    np.testing.assert_allclose(tau.cpu().numpy(), to, rtol=1e-12)
    if Tw is None:
        assert np.all(qw.cpu().numpy() == 0.0)
    else:
        np.testing.assert_allclose(qw.cpu().numpy(), qo, rtol=1e-12, atol=1e-12 * np.abs(qo).max())
    # fallbacks: exactly the points outside the trained log Re_h / T_h/T_w box
    gas = oracle.Gas()
    Uc = np.maximum(U, 1e-8)
    T_wall_eff = Th + gas.Pr ** (1 / 3) * Uc ** 2 / (2 * gas.cp) if Tw is None else np.full_like(Th, Tw)
    rho_w = p / (gas.R * T_wall_eff)
    Reh = h * Uc * rho_w / gas.mu(T_wall_eff)
    lr, tr = np.log(Reh), Th / T_wall_eff
    out = ~((lr >= m.lo[0]) & (lr <= m.hi[0]) & (tr >= m.lo[2]) & (tr <= m.hi[2]))
    near = (np.abs(lr - m.lo[0]) < 1e-9) | (np.abs(lr - m.hi[0]) < 1e-9)
    assert not near.any()
    nfb = int(m.fallbacks.item())
    print(f"[learned-wm] T_wall {Tw}: {nfb} fallbacks of {len(h)} ({out.sum()} out of range)", flush=True)
    assert nfb == int(out.sum())


def test_network_gain_over_the_method_without_it(trained):
    """What the network buys: the Newton steps needed to reach roundoff.  From the
    reference's start clip(sqrt(Re_h), 1, 25) the fixed-count Newton needs many more
    steps than from the network's u+ (and the same small count leaves the reference
    start far from the root)."""
    h, U, Th, p, T_w = _wall_points(1 << 20, 4, out_of_range=False)
    args = [torch.tensor(a, device=DEV) for a in (h, U, Th, p)]
    gas = wmx.Gas()
    to, _ = oracle.EquilibriumWallModel()(oracle.Gas(), h[:20000], U[:20000], Th[:20000], p[:20000], T_w)

    def newton_err(n_iter):
        t, _ = wmx.EquilibriumWallModel(n_iter=n_iter)(gas, *[a[:20000] for a in args], T_w)
        return float(np.max(np.abs(t.cpu().numpy() / to - 1)))

    def learned_err(n_polish):
        # the residual fallback is switched off (tol = inf) so the error is that of the
        # network start plus n_polish Newton steps alone (all points are in range)
        mm = wmx.LearnedWallModel()
        mm.load_state_dict(trained.state_dict())
        mm.n_polish, mm.tol = n_polish, float("inf")
        t, _, _ = mm.eval(gas, *[a[:20000] for a in args], T_wall=T_w)
        return float(np.max(np.abs(t.cpu().numpy() / to - 1)))

    n_newton = next(n for n in range(1, 31) if newton_err(n) <= 1e-12)
    errs = {k: learned_err(k) for k in range(0, 5)}
    n_learned = next(k for k in range(0, 5) if errs[k] <= 1e-12)
    assert errs[0] > 1e-5                         # the raw network alone is not enough
    print(f"[learned-wm] Newton steps to 1e-12: reference start {n_newton}, network start "
          f"{n_learned}; learned-start errors {errs}; 3-step reference-start error "
          f"{newton_err(3):.2e}", flush=True)
    assert n_learned < n_newton
    assert newton_err(n_learned) > 1e-6

    # kernel time on 2^23 points (large enough that launch overheads do not count);
    # reported, not asserted: FP64 Newton steps are cheap on this chip
    big = [a.repeat(8) for a in args]

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    ref30 = wmx.EquilibriumWallModel()
    t_newton = timed(lambda: ref30(gas, *big, T_w))
    t_learned = timed(lambda: trained.eval(gas, *big, T_wall=T_w))
    print(f"[learned-wm] 2^23 points: Newton-30 {t_newton * 1e3:.3f} ms, learned "
          f"{t_learned * 1e3:.3f} ms ({t_newton / t_learned:.2f}x)", flush=True)


def test_learned_model_in_the_solver_equals_newton(trained):
    """The wall kernel with the learned model reproduces the Newton step."""
    spec, arrays = cases.channel_spec(2, 32, 24, 16, {"kind": "KEP", "order": 2}, "wale", T_wall=300.0)
    mesh, s = cases.build(wmx, spec, arrays)
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    R0 = s.rhs(Q)
    st0 = s.wall_stats(Q)
    A = Q
    for _ in range(3):
        A = s.advance(A, s.max_dt(A))
    s.set_wall_net(trained)
    R1 = s.rhs(Q)
    st1 = s.wall_stats(Q)
    B = Q
    for _ in range(3):
        B = s.advance(B, s.max_dt(B))
    for a, b in zip(R0, R1):
        scale = a.abs().amax(dim=(1, 2, 3), keepdim=True)
        assert float(((a - b).abs() / scale).max()) < 1e-12
    for k in st0:
        for x, y in zip(st0[k], st1[k]):
            x, y = np.asarray(torch.as_tensor(x).cpu()), np.asarray(torch.as_tensor(y).cpu())
            np.testing.assert_allclose(y, x, rtol=1e-12, atol=1e-12 * np.abs(x).max())
    for a, b in zip(A, B):
        scale = a.abs().amax(dim=(1, 2, 3), keepdim=True)
        assert float(((a - b).abs() / scale).max()) < 1e-12
    # built with the model on the WallModel faces directly
    spec2, arrays2 = cases.channel_spec(2, 32, 24, 16, {"kind": "KEP", "order": 2}, "wale", T_wall=300.0)
    for e in spec2["bcs"]:
        if e[2]["kind"] == "WallModel":
            e[2]["model"] = trained
    _, s2 = cases.build(wmx, spec2, arrays2)
    R2 = s2.rhs(Q)
    for a, b in zip(R1, R2):
        assert torch.equal(a, b)

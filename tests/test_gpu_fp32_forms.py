"""FP32-bulk form on the device, batched and curvilinear.

* The 16-case uneven channel ensemble (seeds 0-15, Re_b log-uniform in [5e3, 4e4],
  64..192 x 48..96 x 32..64 cells, 2-4 unequal blocks per case) stepped as one plan with
  graph-captured steps equals the same cases stepped one by one, bit for bit, and stays
  within 1e-4 of the FP64 form stepped from the same states on u_tau, the plane-mean
  u(y) and q_w.
* Cases of one ensemble with different mean pressures keep their own FP32 gauge.
* The smooth-ramp block (inflow, sponge, adiabatic wall model) in the FP32-bulk form:
  RHS within 2e-4 of the reference's float64 RHS (golden G5) and of the FP64 form, and
  the spanwise-mean u and wall-mean tau_w within 1e-4 of the FP64 form after steps.

Tolerances: 2e-4 for a single RHS, max-normalised per component (FP32 resolution of
the face fluxes); 1e-4 relative for the mean-flow figures (u_tau, mean u(y)), the
precision form's target; q_w is compared as the wall-mean difference over the largest
FP64 face value, since the isothermal walls start at the bulk temperature and the
mean heat flux is a small difference of face values of both signs.
"""
import numpy as np
import pytest

import cases
import golden_io

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402


def host(Q):
    return [q.detach().cpu().numpy() for q in Q]


def _build(specs, precision):
    solvers, fns = [], []
    for spec, arrays, pr, seed in specs:
        _, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": precision})
        solvers.append(s)
        fns.append(lambda xc, pr=pr, seed=seed: cases.channel_prim(xc, pr, seed))
    return solvers, fns


def _flow_figures(s, Q, rho0):
    """(u_tau from the wall-mean tau_w, plane-mean u(y) over all blocks, per-face q_w)."""
    st = s.wall_stats(Q)
    tau = np.concatenate([v[0].ravel() for v in st.values()])
    qw = np.concatenate([v[1].ravel() for v in st.values()])
    qs = host(Q)
    u = np.concatenate([b.interior(q[1] / q[0]) for b, q in zip(s.blocks, qs)], axis=0)
    return float(np.sqrt(tau.mean() / rho0)), u.mean(axis=(0, 2)), qw


def _compare(fig32, fig64):
    ut32, u32, q32 = fig32
    ut64, u64, q64 = fig64
    return (abs(ut32 - ut64) / ut64, float(np.abs(u32 - u64).max() / np.abs(u64).max()),
            abs(float(q32.mean()) - float(q64.mean())) / float(np.abs(q64).max()))


def test_fp32_ensemble16_equals_cases_and_tracks_fp64():
    specs = cases.ensemble_specs(16)
    print("[ensemble16] building", flush=True)
    s32, fns = _build(specs, "fp32")
    ens32 = wmx.EnsembleSolver(s32)
    assert ens32.n_cases == 16 and sum(len(s.blocks) for s in s32) > 32
    Q0 = ens32.init_state(fns)
    n_steps = 10
    Qe = Q0
    for _ in range(n_steps):                    # captured on the first calls, replayed after
        Qe, dte = ens32.step(Qe, graph=True)
    torch.cuda.synchronize()
    print("[ensemble16] fp32 ensemble stepped", flush=True)
    # every case's own gauge, and its own steps, one case after another
    gauges = ens32._gauge
    assert len(gauges) == 16
    for c, s in enumerate(s32):
        assert s._gauge == gauges[c]
        Q = Q0[c]
        for _ in range(n_steps):
            Q, dt = s.step(Q)
        assert float(dt) == float(dte[c])
        for a, b in zip(Qe[c], Q):
            assert torch.equal(a, b), f"case {c}"
    print("[ensemble16] bitwise equal to per-case steps", flush=True)
    # the FP64 form of the same ensemble from the same states
    s64, _ = _build(specs, "fp64")
    ens64 = wmx.EnsembleSolver(s64)
    Q64 = Q0
    for _ in range(n_steps):
        Q64, _ = ens64.step(Q64)
    torch.cuda.synchronize()
    worst = np.zeros(3)
    for c, (spec, arrays, pr, seed) in enumerate(specs):
        e = _compare(_flow_figures(s32[c], Qe[c], pr["rho0"]), _flow_figures(s64[c], Q64[c], pr["rho0"]))
        print(f"case {c}: u_tau {e[0]:.2e} mean u(y) {e[1]:.2e} q_w {e[2]:.2e}", flush=True)
        worst = np.maximum(worst, e)
        assert e[0] <= 1e-4 and e[1] <= 1e-4 and e[2] <= 1e-4, (c, e)
    print("[ensemble16] worst fp32 vs fp64:", worst)


def test_fp32_ensemble_cases_keep_their_own_gauge():
    """Two cases whose mean pressures differ by 8x: each ensemble case carries its own
    gauge, so both match their per-case FP32 steps bit for bit and their FP64 RHS at FP32
    resolution (a shared gauge would leave the second case's pressure differences at
    the first case's pressure level)."""
    specs = []
    for c, rho0 in enumerate((1.2, 0.15)):
        spec, arrays = cases.channel_spec(2, 32, 16, 16, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                          "vreman", Re_b=1e4, rho0=rho0, name=f"case{c}")
        specs.append((spec, arrays, spec["props"], c))
    assert specs[0][2]["p0"] / specs[1][2]["p0"] == pytest.approx(8.0)
    s32, fns = _build(specs, "fp32")
    s64, _ = _build(specs, "fp64")
    ens = wmx.EnsembleSolver(s32)
    Q0 = ens.init_state(fns)
    Qe = Q0
    for _ in range(3):
        Qe, _ = ens.step(Qe, graph=True)
    g0, g1 = ens._gauge
    assert g0[0] == pytest.approx(specs[0][2]["p0"], rel=1e-3)
    assert g1[0] == pytest.approx(specs[1][2]["p0"], rel=1e-3)
    for c, s in enumerate(s32):
        Q = Q0[c]
        for _ in range(3):
            Q, _ = s.step(Q)
        for a, b in zip(Qe[c], Q):
            assert torch.equal(a, b), c
        R32, R64 = host(s.rhs(Qe[c])), host(s64[c].rhs(Qe[c]))
        for blk, a, b in zip(s.blocks, R32, R64):
            err = golden_io.rel_err(blk.interior(a), blk.interior(b))
            print(f"case {c} p0 {specs[c][2]['p0']:.0f}: fp32 RHS vs fp64 {err:.2e}")
            assert err <= 2e-4, (c, err)


def test_fp32_ramp_rhs_close_to_reference_golden():
    """The smooth-ramp golden case (inflow at two times, sponge, adiabatic wall model,
    far field, slip top) in the FP32-bulk form against the reference's float64 RHS."""
    G5 = golden_io.load("g5_ramp.npz")
    spec, arrays, ent = golden_io.case(G5, "ramp")
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    Q = golden_io.blocks_of(ent, "Qin")
    pr = spec["ramp_props"]
    for it, t in enumerate([0.0, 0.05 * pr["L"] / pr["U_inf"]]):
        R = host(s.rhs(Q, t))
        for blk, r, g in zip(mesh.blocks, R, golden_io.blocks_of(ent, f"R{it}")):
            err = golden_io.rel_err(blk.interior(r), g)
            print("ramp golden fp32", it, err)
            assert err <= 2e-4, (it, err)


def test_fp32_ramp_steps_close_to_fp64():
    """Smooth ramp 48 x 24 x 16 (inflow + sponge + wall model): FP32-bulk against FP64
    for one RHS of a perturbed state, then after 20 steps from the smooth initial state
    the spanwise-mean u(x, y) and the wall-mean tau_w.  (The RHS of the smooth initial
    state itself is a near-cancellation of O(rho U / dx) face fluxes -- the flow is nearly
    steady there -- so its max-normalised difference measures the FP32 face-flux rounding
    against a vanishing RHS, not the form's accuracy; the perturbed state has the RHS
    magnitude of a developed flow.)"""
    spec, arrays, prim = cases.ramp_spec(wmx, 48, 24, 16, n_modes=16)
    out = {}
    for prec in ("fp64", "fp32"):
        mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": prec})
        Q0 = s.init_state(prim)
        R = host(s.rhs(cases.interior_noise(mesh.blocks, host(Q0), 7, 0.02), 0.0))
        Q, t = Q0, 0.0
        for _ in range(20):
            Qn, dt = s.step(Q, time=t)
            t += float(dt)
            Q = Qn
        st = s.wall_stats(Q)
        out[prec] = (R, host(Q), np.mean([v[0].mean() for v in st.values()]), t, mesh)
    blk = out["fp64"][4].blocks[0]
    err = golden_io.rel_err(blk.interior(out["fp32"][0][0]), blk.interior(out["fp64"][0][0]))
    print("ramp rhs fp32 vs fp64", err)
    assert err <= 2e-4
    q64, q32 = out["fp64"][1][0], out["fp32"][1][0]
    u64 = blk.interior(q64[1] / q64[0]).mean(axis=2)
    u32 = blk.interior(q32[1] / q32[0]).mean(axis=2)
    eu = float(np.abs(u32 - u64).max() / np.abs(u64).max())
    et = abs(out["fp32"][2] - out["fp64"][2]) / abs(out["fp64"][2])
    print(f"ramp after 20 steps: mean u {eu:.2e}, tau_w {et:.2e}, t {out['fp64'][3]:.3e} s")
    assert eu <= 1e-4 and et <= 1e-4


@pytest.mark.parametrize("kind", ["channel", "rot_cyc", "ramp"])
def test_fused_gradients_match_the_separate_gradient_kernel(kind):
    """The fused FP32-bulk kernel (gradient records evaluated in the flux march) against the
    separate gradient kernel + flux kernel: the same FP32 arithmetic per gradient record, so
    the RHS agrees to FP32 rounding of the records (bar 2e-6, max-normalised per component;
    measured 4e-7) and one step to 2e-8 relative (the RHS difference times dt R / q)."""
    if kind == "channel":
        spec, arrays = cases.channel_spec(2, 48, 32, 16, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                          "vreman", Re_b=2e4, splits=[0, 20, 48])
        pr = spec["props"]
        prim = lambda xc: cases.channel_prim(xc, pr, 2)
    elif kind == "rot_cyc":
        spec, arrays = cases.multiblock_spec("rot_cyc", (12, 10, 8), {"kind": "Hybrid"},
                                             {"kind": "WallModel", "T_wall": 300.0, "n_match": 1}, sgs="wale")
        prim = cases.prim_fn
    else:
        spec, arrays, prim = cases.ramp_spec(wmx, 48, 24, 16, n_modes=8)
    out = {}
    for fused in (True, False):
        mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
        s.set_fused_gradients(fused)
        Q0 = s.init_state(prim)
        Qn = cases.interior_noise(mesh.blocks, host(Q0), 11, 0.02)
        R = host(s.rhs(Qn, 0.0))
        Q1, dt = s.step(Qn, time=0.0)
        out[fused] = (R, host(Q1), float(dt), mesh)
    mesh = out[True][3]
    assert out[True][2] == out[False][2]
    for blk, a, b in zip(mesh.blocks, out[True][0], out[False][0]):
        err = golden_io.rel_err(blk.interior(a), blk.interior(b))
        print(kind, "fused vs separate RHS", err)
        assert err <= 2e-6, err
    for a, b in zip(out[True][1], out[False][1]):
        err = float(np.abs(a - b).max() / np.abs(b).max())
        print(kind, "fused vs separate step", err)
        assert err <= 2e-8, err

"""Batching form on the device: graph-captured steps and merged-case ensembles must
reproduce the per-case, launch-by-launch step bit for bit; the device plane statistics
against the oracle."""
import numpy as np
import pytest

import cases
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402


def _channel(nblocks, nx, ny, nz, Re_b, seed, splits=None, precision="fp64"):
    spec, arrays = cases.channel_spec(nblocks, nx, ny, nz, {"kind": "KEP", "order": 2}, "wale",
                                      Re_b=Re_b, splits=splits)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": precision})
    pr = spec["props"]
    return s, (lambda xc, pr=pr, seed=seed: cases.channel_prim(xc, pr, seed))


@pytest.mark.parametrize("precision,own_stream", [("fp64", False), ("fp32", False), ("fp32", True)])
def test_graph_step_equals_max_dt_plus_advance(precision, own_stream):
    # own_stream: the caller runs on a stream of its own, which the step captures and
    # replays on directly (on the legacy default stream it uses a side stream)
    s, fn = _channel(2, 32, 16, 16, 2e4, 0, precision=precision)
    Q0 = s.init_state(fn)
    A, B = Q0, Q0
    st = torch.cuda.Stream() if own_stream else torch.cuda.current_stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        for _ in range(4):                   # graph captured on the 1st/2nd call, replayed after
            A, dta = s.step(A, graph=True)
            dtb = s.max_dt(B)
            B = s.advance(B, dtb)
        torch.cuda.synchronize()
    assert float(dta) == float(dtb)
    for a, b in zip(A, B):
        assert torch.equal(a, b)


def test_ensemble_equals_individual_cases():
    specs = [(2, 32, 16, 16, 2e4, 0, None), (3, 48, 16, 8, 8e3, 1, [0, 10, 30, 48]),
             (2, 24, 24, 16, 3e4, 2, [0, 9, 24])]
    solvers, fns = [], []
    for nb, nx, ny, nz, re, seed, sp in specs:
        s, fn = _channel(nb, nx, ny, nz, re, seed, sp)
        solvers.append(s)
        fns.append(fn)
    ens = wmx.EnsembleSolver(solvers)
    Qs = ens.init_state(fns)
    ref = [Q for Q in Qs]
    side = torch.cuda.Stream()
    for it in range(3):
        if it == 2:                          # the last step replayed on a stream of the caller's own
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                Qs, dts = ens.step(Qs, graph=True)
            torch.cuda.current_stream().wait_stream(side)
        else:
            Qs, dts = ens.step(Qs, graph=True)
        new = []
        for c, (s, Q) in enumerate(zip(solvers, ref)):
            dt = s.max_dt(Q)
            assert float(dt) == float(dts[c])
            new.append(s.advance(Q, dt))
        ref = new
    torch.cuda.synchronize()
    for Qe, Qr in zip(Qs, ref):
        for a, b in zip(Qe, Qr):
            assert torch.equal(a, b)


def test_autotuner_picks_a_candidate(tmp_path, monkeypatch):
    monkeypatch.setenv("WMX_TUNE_CACHE", str(tmp_path / "t.json"))
    s, fn = _channel(1, 48, 32, 32, 2e4, 0)
    Q = s.init_state(fn)
    best, timings = wmx.autotune.tune(s, Q, steps=1, candidates=((8, 0), (16, 1)))
    assert best in ((8, 0), (16, 1)) and len(timings) == 2
    best2, _ = wmx.autotune.tune(s, Q)          # cached
    assert best2 == best
    A, _ = s.step(Q, graph=False)
    s.set_tiling(16, 0)
    B, _ = s.step(Q, graph=False)
    for a, b in zip(A, B):                       # tiling changes scheduling, not results
        assert torch.equal(a, b)


def test_plane_stats_match_the_oracle():
    """PlaneStats on the device (wmx_plane_stats) against the oracle's plane means over two
    accumulated states of a channel block, the Reynolds stresses and the wall units."""
    from oracle import stats as OS
    spec, arrays = cases.channel_spec(1, 24, 20, 16, {"kind": "KEP", "order": 2}, "wale")
    mesh, s = cases.build(wmx, spec, arrays)
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    st = wmx.PlaneStats(mesh, s)
    blk = mesh.blocks[0]
    ost = OS.PlaneStats(st.y)
    gas = oracle.Gas()
    for _ in range(2):
        Q = s.advance(Q, s.max_dt(Q))
        st.accumulate(Q)
        q = blk.interior(Q[0].cpu().numpy())
        tau = np.mean([np.mean(v[0]) for v in s.wall_stats(Q).values()])
        ost.accumulate(q, gas, tau)
    y, avg = st.result()
    _, ref = ost.result()
    for k in ref:
        scale = max(np.abs(ref[k]).max(), np.abs(ost.sum[k]).max() / ost.n * 1e-3, 1e-300)
        assert np.abs(avg[k] - ref[k]).max() <= 1e-12 * scale, k
    assert abs(st.mean_tau_w - ost.mean_tau_w) <= 1e-14 * abs(ost.mean_tau_w)
    rho_w, mu_w = pr["rho0"], pr["mu0"]
    ut, out = st.wall_units(st.mean_tau_w, rho_w, mu_w, pr["d"])
    ut_o, out_o = ost.wall_units(ost.mean_tau_w, rho_w, mu_w, pr["d"])
    assert abs(ut - ut_o) <= 1e-14 * ut_o
    for k in out_o:
        np.testing.assert_allclose(out[k], out_o[k], rtol=1e-10, atol=1e-12 * np.abs(out_o[k]).max())


def test_blocks_interleaved_across_cases():
    """The same two-case plan with its blocks interleaved across the cases (per-case
    workgroup lists that are not the identity: the general path of the per-case
    reductions) gives the per-case dt and the mass-flux constraint bit for bit."""
    import ctypes as C
    from wmles_jax_amd import _lib
    specs = [(2, 32, 16, 16, 2e4, 0, None), (3, 48, 16, 8, 8e3, 1, [0, 10, 30, 48])]
    solvers, fns = [], []
    for nb, nx, ny, nz, re, seed, sp in specs:
        s, fn = _channel(nb, nx, ny, nz, re, seed, sp)
        solvers.append(s)
        fns.append(fn)
    ens = wmx.EnsembleSolver(solvers)
    Qs = ens.init_state(fns)
    flat = ens._flat(Qs)
    nblk = len(flat)
    assert nblk == 5
    perm = [0, 2, 1, 3, 4]                     # plan position -> case-ordered block
    inv = {old: new for new, old in enumerate(perm)}
    descs = (_lib.BlockDesc * nblk)()
    for new, old in enumerate(perm):
        d = _lib.BlockDesc()
        C.memmove(C.byref(d), C.byref(ens._descs[old]), C.sizeof(_lib.BlockDesc))
        for f in range(6):
            if d.faces[f].kind == _lib.FACE_INTERFACE:
                d.faces[f].nb = inv[d.faces[f].nb]
        descs[new] = d
    plan = C.c_void_p()
    _lib.check(_lib.lib.wmx_plan_create(nblk, descs, C.byref(ens._nd), 0, None, C.byref(plan)))
    try:
        nbytes = C.c_size_t()
        _lib.check(_lib.lib.wmx_plan_workspace_bytes(plan, C.byref(nbytes)))
        ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=flat[0].device)
        _lib.check(_lib.lib.wmx_plan_bind_workspace(plan, C.c_void_p(ws.data_ptr()), nbytes.value))
        dt_ref = ens.max_dt(Qs)
        out = torch.empty(16, dtype=torch.float64, device=flat[0].device)
        _lib.check(_lib.lib.wmx_max_dt(plan, _lib.ptr_array([flat[o] for o in perm]),
                                       C.c_void_p(out.data_ptr()), _lib.stream_ptr()))
        torch.cuda.synchronize()
        assert torch.equal(out[:2], dt_ref)
        A = [q.clone() for q in flat]
        B = [q.clone() for q in flat]
        _lib.check(_lib.lib.wmx_mass_flux(ens._plan, _lib.ptr_array(A), _lib.stream_ptr()))
        _lib.check(_lib.lib.wmx_mass_flux(plan, _lib.ptr_array([B[o] for o in perm]), _lib.stream_ptr()))
        torch.cuda.synchronize()
        assert any(not torch.equal(a, q) for a, q in zip(A, flat))     # the constraint acted
        for a, b in zip(A, B):
            assert torch.equal(a, b)
    finally:
        torch.cuda.synchronize()
        _lib.lib.wmx_plan_destroy(plan)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_step_with_inflow_equals_max_dt_plus_advance(precision):
    """Time-dependent inflow: Solver.step takes the separate max_dt + advance path (the
    fill reads dt), and its result equals those two calls bit for bit."""
    spec, arrays, prim = cases.ramp_spec(wmx, 48, 24, 16, n_modes=8)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": precision})
    Q0 = s.init_state(prim)
    for t in (0.0, 1.5e-6):
        A, dta = s.step(Q0, time=t, graph=False)
        dtb = s.max_dt(Q0, t)
        B = s.advance(Q0, dtb, t)
        torch.cuda.synchronize()
        assert float(dta) == float(dtb)
        for a, b in zip(A, B):
            assert torch.equal(a, b)


@pytest.mark.parametrize("field", ["n_tab", "n_modes"])
def test_plan_rejects_an_empty_inflow_table(field):
    """An inflow descriptor with no profile rows or no modes is refused by wmx_plan_create
    with a message (the fill's table search needs a row, its phase sum a mode)."""
    import ctypes as C
    from wmles_jax_amd import _lib
    spec, arrays, prim = cases.ramp_spec(wmx, 48, 24, 16, n_modes=8)
    mesh, s = cases.build(wmx, spec, arrays)
    inf = s._ctx.inflows[0]
    descs = (_lib.InflowDesc * 1)()
    descs[0] = inf.desc(s._dev)
    setattr(descs[0], field, 0)
    plan = C.c_void_p()
    with pytest.raises(_lib.WmxError, match="profile table and at least one mode"):
        _lib.check(_lib.lib.wmx_plan_create(len(s._descs), s._descs, C.byref(s._nd), 1, descs,
                                            C.byref(plan)))
    assert not plan.value

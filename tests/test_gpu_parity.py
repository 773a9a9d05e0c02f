"""HIP step against the reference golden vectors and the CPU oracle (MI355X).

Bars: interface ghost maps bit-exact; BC ghost fills and geometry bit-exact (their
units are built without FP contraction); RHS, advance, wall stats within 1e-12 of
the reference/oracle, measured as max|diff| / max|ref| per component (the
reference's own normalisation, tests/test_multiblock.py:94-95).
"""
import numpy as np
import pytest

import cases
import golden_io
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402

RTOL = 1e-12

G1 = golden_io.load("g1_multiblock.npz")
G2 = golden_io.load("g2_unique_ids.npz")
G5 = golden_io.load("g5_ramp.npz")


def host(Q):
    return [q.detach().cpu().numpy() for q in Q]


def _times(spec, ent):
    if "ramp_props" in spec:
        pr = spec["ramp_props"]
        return [0.0, 0.05 * pr["L"] / pr["U_inf"]]
    return [0.0]


def _check_golden(npz, key):
    spec, arrays, ent = golden_io.case(npz, key)
    print(f"[golden] {key}", flush=True)
    mesh, s = cases.build(wmx, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    times = _times(spec, ent)
    for it, t in enumerate(times):
        Qbc = host(s.apply_bcs(Q, s.geoms, t))
        for q, g in zip(Qbc, golden_io.blocks_of(ent, f"Qbc{it}")):
            err = golden_io.rel_err(q, g)
            inflow = "ramp_props" in spec
            assert err <= (1e-13 if inflow else 0.0), (key, it, err)
    for it, t in enumerate(times):
        R = host(s.rhs(Q, t))
        for blk, r, g in zip(mesh.blocks, R, golden_io.blocks_of(ent, f"R{it}")):
            err = golden_io.rel_err(blk.interior(r), g)
            assert err <= RTOL, (key, "rhs", it, err)
            ng = blk.ng
            assert np.all(r[:, :ng] == 0.0) and np.all(r[:, -ng:] == 0.0)
    dt = float(s.max_dt(Q))
    assert abs(dt - float(ent["dt"])) <= 1e-14 * float(ent["dt"]), (key, dt, float(ent["dt"]))
    Qa = host(s.advance(Q, float(ent["dt"]), times[-1]))
    for q, g in zip(Qa, golden_io.blocks_of(ent, "Qadv")):
        err = golden_io.rel_err(q, g)
        assert err <= RTOL, (key, "advance", err)
    for (bname, face), (tau, qw) in s.wall_stats(Q).items():
        np.testing.assert_allclose(tau, ent[f"tau_{bname}_{face}"], rtol=1e-12)
        gq = ent[f"qw_{bname}_{face}"]
        np.testing.assert_allclose(qw, gq, rtol=1e-11, atol=1e-12 * max(1.0, np.abs(gq).max()))


@pytest.mark.parametrize("key", golden_io.case_keys(G1))
def test_golden_multiblock(key):
    _check_golden(G1, key)


def test_golden_ramp_inflow_sponge():
    _check_golden(G5, "ramp")


@pytest.mark.parametrize("key", golden_io.case_keys(G2))
def test_unique_id_ghost_maps_bit_exact(key):
    spec, arrays, ent = golden_io.case(G2, key)
    print(f"[ids] {key}", flush=True)
    mesh, s = cases.build(wmx, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    for q, g in zip(host(s.apply_bcs(Q)), golden_io.blocks_of(ent, "Qbc0")):
        np.testing.assert_array_equal(q, g)


@pytest.mark.parametrize("kind", ["single", "split", "flip", "rot_cyc"])
def test_geometry_bit_exact(kind):
    spec, arrays = cases.multiblock_spec(kind, (12, 10, 8))
    m0, _ = cases.build(oracle, spec, arrays)
    m1, _ = cases.build(wmx, spec, arrays)
    for g0, g1 in zip(m0.geoms, m1.geoms):
        np.testing.assert_array_equal(g1.vol.cpu().numpy(), g0.vol)
        np.testing.assert_array_equal(g1.xc.cpu().numpy(), g0.xc)
        for d in range(3):
            np.testing.assert_array_equal(g1.S[d].cpu().numpy(), g0.S[d])
            np.testing.assert_array_equal(g1.xf[d].cpu().numpy(), g0.xf[d])


def test_channel_two_steps_golden():
    npz = golden_io.load("g4_channel.npz")
    spec, arrays, ent = golden_io.case(npz, "channel")
    mesh, s = cases.build(wmx, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    t = 0.0
    for step in range(2):
        dt = s.max_dt(Q)
        assert abs(float(dt) - float(ent[f"dt{step}"])) <= 1e-13 * float(dt)
        Q = s.advance(Q, dt, t)                 # device dt, no host round trip
        t += float(dt)
        for q, g in zip(host(Q), golden_io.blocks_of(ent, f"Q{step}")):
            assert golden_io.rel_err(q, g) <= RTOL
    for (bname, face), (tau, qw) in s.wall_stats(Q).items():
        np.testing.assert_allclose(tau, ent[f"tau_{bname}_{face}"], rtol=1e-11)
        np.testing.assert_allclose(qw, ent[f"qw_{bname}_{face}"], rtol=1e-10)


def _oracle_vs_hip(spec, arrays, state_fn, steps=1, rtol=RTOL):
    m0, s0 = cases.build(oracle, spec, arrays)
    m1, s1 = cases.build(wmx, spec, arrays)
    Q = state_fn(m0, s0)
    R0, R1 = s0.rhs(Q), host(s1.rhs(Q))
    for blk, a, b in zip(m0.blocks, R1, R0):
        err = golden_io.rel_err(blk.interior(a), blk.interior(b))
        assert err <= rtol, ("rhs", err)
    Qo, Qh = Q, Q
    for _ in range(steps):
        dt = float(s0.max_dt(Qo))
        dth = float(s1.max_dt(Qh))
        assert abs(dt - dth) <= 1e-13 * dt
        Qo = s0.advance(Qo, dt)
        Qh = s1.advance(Qh, dt)
    for a, b in zip(host(Qh), Qo):
        err = golden_io.rel_err(a, b)
        assert err <= rtol * 10 ** (steps - 1), ("advance", err)
    return m1, s1


def test_channel_config1_vs_oracle():
    """Plain-form channel: 2 blocks x 32x32x16, KEP(2), WALE, wall model, mass flux."""
    spec, arrays = cases.channel_spec(2, 64, 32, 16, {"kind": "KEP", "order": 2}, "wale")
    pr = spec["props"]

    def state(m, s):
        return s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    _oracle_vs_hip(spec, arrays, state, steps=2)


def test_channel_hybrid_vreman_vs_oracle():
    """The cfg2 numerics (Hybrid KEP4 + MUSCL-HLLC, Vreman, isothermal WM) on a small
    single-block channel with noise strong enough to switch the hybrid on."""
    spec, arrays = cases.channel_spec(1, 24, 16, 16, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman")
    pr = spec["props"]

    def state(m, s):
        return s.init_state(lambda xc: cases.channel_prim(xc, pr, 3, noise=0.2))
    _oracle_vs_hip(spec, arrays, state, steps=1)


@pytest.mark.parametrize("scheme", [{"kind": "KEP", "order": 2}, {"kind": "KEP", "order": 4},
                                    {"kind": "KEEP", "order": 4}, {"kind": "MUSCLHLLC"},
                                    {"kind": "Hybrid"}])
def test_freestream_preservation(scheme):
    spec, arrays = cases.multiblock_spec("single", (12, 10, 8), scheme, periodic_y=True)
    mesh, s = cases.build(wmx, spec, arrays)
    Q = s.init_state(lambda xc: np.stack([1.2 + 0 * xc[0], 50 + 0 * xc[0], -20 + 0 * xc[0],
                                          30 + 0 * xc[0], 1e5 + 0 * xc[0]]))
    R = host(s.rhs(Q))[0]
    assert np.abs(mesh.blocks[0].interior(R)).max() < 1e-6


@pytest.mark.parametrize("kind", ["split", "flip", "rot_cyc"])
def test_multiblock_equals_single_block(kind):
    def gather(m, R):
        out = {}
        for blk, r, g in zip(m.blocks, R, m.geoms):
            xc = blk.interior(g.xc.cpu().numpy()).reshape(3, -1).T
            rr = blk.interior(r).reshape(5, -1).T
            for c, v in zip(xc, rr):
                out[tuple(np.round(c, 8))] = v
        return out
    res = {}
    for k in ("single", kind):
        spec, arrays = cases.multiblock_spec(k, (12, 10, 8), {"kind": "Hybrid"},
                                             {"kind": "WallModel", "T_wall": 300.0, "n_match": 1})
        m, s = cases.build(wmx, spec, arrays)
        res[k] = gather(m, host(s.rhs(s.init_state(cases.prim_fn))))
    R0, R1 = res["single"], res[kind]
    assert R0.keys() == R1.keys()
    scale = np.max(np.abs(np.array(list(R0.values()))))
    assert max(np.max(np.abs(R0[k] - R1[k])) for k in R0) / scale < 1e-12


def test_characteristic_farfield_known_answers():
    """tests/test_multiblock.py:110-131 through BC.fill on the device."""
    spec, arrays = cases.multiblock_spec("single", (12, 10, 8))
    mesh, s = cases.build(wmx, spec, arrays)
    blk, geom = mesh.blocks[0], mesh.geoms[0]
    ng, n = blk.ng, blk.n[1]
    prim_inf = np.array([1.2, 10.0, 0.0, 0.0, 1e5])
    bc = wmx.CharacteristicFarfield(prim_inf)

    def state(prim):
        w = np.asarray(prim, float).reshape(5, 1, 1, 1) * np.ones((1,) + blk.N)
        return torch.as_tensor(oracle.Gas().prim_to_cons(w), device="cuda")
    for q, side, expected in [(state([1.0, 0.0, 500.0, 0.0, 8e4]), 0, prim_inf),
                              (state([1.0, 0.0, 1000.0, 0.0, 8e4]), 1, [1.0, 0.0, 1000.0, 0.0, 8e4]),
                              (state(prim_inf), 0, prim_inf), (state(prim_inf), 1, prim_inf)]:
        filled = bc.fill(q, blk, geom, s.gas, 1, side).cpu().numpy()
        gj = ng - 1 if side == 0 else ng + n
        got = oracle.Gas().cons_to_prim(filled[:, ng, gj, ng])[0]
        np.testing.assert_allclose(got, expected, rtol=1e-12)


@pytest.mark.parametrize("ut,nu,h", [(2.0, 1.5e-5, 0.003), (0.5, 1.5e-5, 1e-4), (5.0, 1e-5, 0.01)])
def test_spalding_newton_matches_bisection(ut, nu, h):
    from test_oracle_invariants import spalding_bisection
    U = spalding_bisection(ut, nu, h)
    got = wmx.EquilibriumWallModel().utau(np.array([h]), np.array([U]), np.array([1.2]),
                                          np.array([nu * 1.2]))
    assert abs(float(got[0]) - ut) < 1e-6 * max(1.0, ut)


def test_inputs_unchanged_and_device_dt_matches_host_dt():
    spec, arrays = cases.multiblock_spec("flip", (12, 10, 8), {"kind": "Hybrid"},
                                         {"kind": "WallModel", "T_wall": 300.0, "n_match": 1},
                                         constraints=[{"kind": "ConstantMassFlux", "direction": 0,
                                                       "target": 25.0}])
    mesh, s = cases.build(wmx, spec, arrays)
    Q = s.init_state(cases.prim_fn)
    keep = [q.clone() for q in Q]
    dt = s.max_dt(Q)
    A = s.advance(Q, dt)
    B = s.advance(Q, float(dt))
    s.rhs(Q)
    s.apply_bcs(Q)
    for q, k in zip(Q, keep):
        assert torch.equal(q, k)
    for a, b in zip(A, B):
        assert torch.equal(a, b)


def test_sod_shock_tube_on_device():
    """Sod tube (tests/test_sod.py) through the HIP MUSCL-HLLC step: plateau values
    against the exact Riemann solution and mass conservation."""
    gas = wmx.Gas(R=1.0, mu_ref=1e-9)
    x = np.linspace(0, 1, 101)
    y = np.linspace(0, 0.1, 4)
    b = wmx.Block("tube", np.stack(np.meshgrid(x, y, y, indexing="ij")))
    m = wmx.Mesh([b])
    b.set_bc("imin", wmx.Extrapolate())
    b.set_bc("imax", wmx.Extrapolate())
    m.connect(b, "jmin", b, "jmax", shift=(0, -0.1, 0))
    m.connect(b, "kmin", b, "kmax", shift=(0, 0, -0.1))
    m.finalize()
    s = wmx.Solver(m, gas, wmx.MUSCLHLLC(), sgs=None, cfl=0.5)
    Q = s.init_state(lambda xc: np.stack([np.where(xc[0] < 0.5, 1.0, 0.125), 0 * xc[0], 0 * xc[0],
                                          0 * xc[0], np.where(xc[0] < 0.5, 1.0, 0.1)]))
    vol = b.interior(m.geoms[0].vol)
    mass0 = float((b.interior(Q[0][0]) * vol).sum())
    t = 0.0
    while t < 0.2 - 1e-12:
        dt = min(float(s.max_dt(Q)), 0.2 - t)
        Q = s.advance(Q, dt)
        t += dt
    w, _ = gas.cons_to_prim(Q[0].cpu().numpy())
    w = b.interior(w)[:, :, 1, 1]
    xc = b.interior(m.geoms[0].xc.cpu().numpy())[0, :, 1, 1]
    i, j = np.argmin(abs(xc - 0.75)), np.argmin(abs(xc - 0.60))
    assert abs(w[4, i] - 0.30313) < 0.01 and abs(w[1, i] - 0.92745) < 0.03
    assert abs(w[0, i] - 0.26557) < 0.01 and abs(w[0, j] - 0.42632) < 0.02
    assert abs(float((b.interior(Q[0][0]) * vol).sum()) - mass0) < 1e-12


# ----------------------------------------------------------------------- FP32-bulk form
@pytest.mark.parametrize("key", ["split_hybrid_slip", "rotcyc_hybrid_wm", "single_muscl_noslip",
                                 "flip_kep2_wale", "single_keep4_wmadi"])
def test_fp32_bulk_rhs_close_to_reference(key):
    """FP32-bulk RHS against the reference's float64 RHS: FP32 arithmetic with the
    FP64 gauge terms keeps the per-component error near FP32 resolution."""
    spec, arrays, ent = golden_io.case(G1, key)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    Q = golden_io.blocks_of(ent, "Qin")
    R = host(s.rhs(Q))
    for blk, r, g in zip(mesh.blocks, R, golden_io.blocks_of(ent, "R0")):
        err = golden_io.rel_err(blk.interior(r), g)
        print(key, err)
        assert err <= 2e-4, (key, err)


def test_fp32_bulk_channel_steps_close_to_fp64():
    """Channel (cfg2 numerics, 32x32x16 cells) after 20 steps: FP32-bulk vs FP64 state,
    u_tau and q_w."""
    spec, arrays = cases.channel_spec(1, 32, 32, 16, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman")
    pr = spec["props"]
    out = {}
    for prec in ("fp64", "fp32"):
        mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": prec})
        Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
        for _ in range(20):
            Q = s.advance(Q, s.max_dt(Q))
        out[prec] = (host(Q)[0], s.wall_stats(Q))
    q64, st64 = out["fp64"]
    q32, st32 = out["fp32"]
    blk = mesh.blocks[0]
    u64 = blk.interior(q64[1] / q64[0]).mean(axis=(0, 2))
    u32 = blk.interior(q32[1] / q32[0]).mean(axis=(0, 2))
    assert np.abs(u32 - u64).max() / np.abs(u64).max() < 1e-4
    for key in st64:
        t64, t32 = st64[key][0].mean(), st32[key][0].mean()
        assert abs(t32 - t64) / abs(t64) < 1e-4
        q_64, q_32 = st64[key][1].mean(), st32[key][1].mean()
        assert abs(q_32 - q_64) / max(abs(q_64), 1e-30) < 1e-3


def test_fp32_bulk_wale_where_the_float_expression_overflows():
    """WALE in the FP32-bulk form at millimetre scale: channel at Re_b 5e3 (half height
    0.7 mm) with 64 wall-normal cells.  The WALE numerator (Sd:Sd)^(3/2) of the initial
    state exceeds the float range (checked on the host from finite differences of the
    cell-centre velocities), yet the FP32-bulk RHS stays finite and within FP32
    resolution of the FP64 RHS, and steps stay finite."""
    spec, arrays = cases.channel_spec(1, 32, 64, 16, {"kind": "KEP", "order": 2}, "wale", Re_b=5e3)
    pr = spec["props"]
    X = arrays["nodes_B0"]
    xc = 0.125 * sum(X[:, a:X.shape[1] - 1 + a, b:X.shape[2] - 1 + b, c:X.shape[3] - 1 + c]
                     for a in (0, 1) for b in (0, 1) for c in (0, 1))
    vel = cases.channel_prim(xc, pr, 0)[1:4]
    axes = (xc[0][:, 0, 0], xc[1][0, :, 0], xc[2][0, 0, :])
    g = np.stack([np.stack([np.gradient(vel[i], axes[j], axis=j) for j in range(3)]) for i in range(3)])
    g2 = np.einsum("ik...,kj...->ij...", g, g)
    sd = 0.5 * (g2 + g2.transpose(1, 0, 2, 3, 4)) - np.eye(3).reshape(3, 3, 1, 1, 1) * np.trace(g2) / 3.0
    assert float(((sd * sd).sum(axis=(0, 1)) ** 1.5).max()) > 10 * float(np.finfo(np.float32).max)
    R = {}
    for prec in ("fp64", "fp32"):
        mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": prec})
        Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
        R[prec] = host(s.rhs(Q))
        if prec == "fp32":
            for _ in range(3):
                Q = s.advance(Q, s.max_dt(Q))
            assert all(bool(torch.isfinite(q).all()) for q in Q)
    blk = mesh.blocks[0]
    assert np.isfinite(blk.interior(R["fp32"][0])).all()
    err = golden_io.rel_err(blk.interior(R["fp32"][0]), blk.interior(R["fp64"][0]))
    print("wale fp32 at mm scale", err)
    assert err <= 2e-4


def test_redundant_sweep_passes_dropped_only_when_exact():
    """The plan drops the second sweep pass for the channel (it reproduces pass 1 bit for
    bit) and keeps it where edge ghosts depend on it; golden ghost maps stay exact."""
    spec, arrays = cases.channel_spec(2, 24, 16, 8, {"kind": "KEP", "order": 2}, "wale")
    mesh, s = cases.build(wmx, spec, arrays)
    assert s.fill_phases()[1] == 1
    kept = {}
    for key in golden_io.case_keys(G2):
        spec, arrays, ent = golden_io.case(G2, key)
        mesh, s = cases.build(wmx, spec, arrays)
        kept[key] = s.fill_phases()[1]
        for q, g in zip(host(s.apply_bcs(golden_io.blocks_of(ent, "Qin"))),
                        golden_io.blocks_of(ent, "Qbc0")):
            np.testing.assert_array_equal(q, g)
    print("sweep passes kept:", kept)


def test_inflow_profile_lookup_exact_nodes_and_clamps():
    """The inflow's profile lookup (numpy.interp semantics: one bracket search of the y
    column, shared by the six value columns) against the oracle, with a table whose y
    column holds ghost-cell centroids exactly (462 of the 1980 ghost cells sit on a node)
    and spans less than the ghosted rows (cells below the first and above the last
    node); ghost values at the golden ramp's 1e-13 bar."""
    spec, arrays, prim = cases.ramp_spec(wmx, 48, 24, 16, n_modes=8)
    L, ng = spec["ramp_props"]["L"], spec["ng"]
    m0, _ = cases.build(oracle, spec, arrays)
    yl = np.asarray(m0.geoms[0].xc)[1, :ng] / L          # the inflow ghost cells' y / L
    yt = np.unique(np.concatenate([yl[:, ng + 2:-ng - 2:3, :].ravel(), np.linspace(0.25, 0.7, 9)]))
    assert np.isin(yl, yt).sum() > 0 and yl.min() < yt[0] and yl.max() > yt[-1]
    a = 0.08 ** 2
    prof = np.stack([yt, 0.5 + 0.5 * np.tanh((yt - 0.3) / 0.05), 0.01 * np.sin(9.0 * yt),
                     a * (1.0 + 0.3 * np.sin(20.0 * yt)), a * (1.0 + 0.2 * np.cos(15.0 * yt)),
                     0.8 * a + 0.0 * yt, -0.3 * a * (1.0 + 0.1 * yt)], axis=1)
    arrays = dict(arrays, inflow_profile=prof)
    _, s0 = cases.build(oracle, spec, arrays)
    _, s1 = cases.build(wmx, spec, arrays)
    Q = s1.init_state(prim)
    t = 0.05 * L / spec["ramp_props"]["U_inf"]
    A = host(s1.apply_bcs(Q, s1.geoms, t))
    B = s0.apply_bcs(host(Q), s0.geoms, t)
    for qa, qb in zip(A, B):
        assert golden_io.rel_err(qa[:, :ng], np.asarray(qb)[:, :ng]) <= 1e-13

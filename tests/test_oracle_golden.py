"""The oracle against the reference's own outputs (golden vectors, CPU)."""
import numpy as np
import pytest

import cases
import golden_io
import oracle

G1 = golden_io.load("g1_multiblock.npz")
G2 = golden_io.load("g2_unique_ids.npz")
G5 = golden_io.load("g5_ramp.npz")


def _times(entries):
    ts, i = [], 0
    while f"Qbc{i}_0" in entries:
        ts.append(i)
        i += 1
    return ts


def _check_case(npz, key):
    spec, arrays, ent = golden_io.case(npz, key)
    mesh, solver = cases.build(oracle, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    times = [0.0] if "ramp_props" not in spec else None
    if times is None:
        pr = spec["ramp_props"]
        times = [0.0, 0.05 * pr["L"] / pr["U_inf"]]
    for it in _times(ent):
        Qbc = solver.apply_bcs(Q, solver.geoms, times[it])
        for q, g in zip(Qbc, golden_io.blocks_of(ent, f"Qbc{it}")):
            assert golden_io.rel_err(q, g) <= 1e-15, key
    if "R0_0" not in ent:
        return
    for it in _times(ent):
        R = solver.rhs(Q, times[it])
        for blk, r, g in zip(mesh.blocks, R, golden_io.blocks_of(ent, f"R{it}")):
            assert golden_io.rel_err(blk.interior(r), g) <= 1e-13, key
    dt = float(solver.max_dt(Q))
    assert abs(dt - float(ent["dt"])) <= 1e-14 * abs(float(ent["dt"]))
    Qa = solver.advance(Q, float(ent["dt"]), times[-1])
    for q, g in zip(Qa, golden_io.blocks_of(ent, "Qadv")):
        assert golden_io.rel_err(q, g) <= 1e-13, key
    st = solver.wall_stats(Q)
    for (bname, face), (tau, qw) in st.items():
        gt, gq = ent[f"tau_{bname}_{face}"], ent[f"qw_{bname}_{face}"]
        np.testing.assert_allclose(tau, gt, rtol=1e-12, atol=0)
        np.testing.assert_allclose(qw, gq, rtol=1e-11, atol=1e-12 * max(1.0, np.abs(gq).max()))


@pytest.mark.parametrize("key", golden_io.case_keys(G1))
def test_multiblock_golden(key):
    _check_case(G1, key)


@pytest.mark.parametrize("key", golden_io.case_keys(G2))
def test_unique_id_ghost_maps(key):
    """Interface copies of unique ids reproduce the reference sweep bit for bit."""
    spec, arrays, ent = golden_io.case(G2, key)
    mesh, solver = cases.build(oracle, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    Qbc = solver.apply_bcs(Q, solver.geoms, 0.0)
    for q, g in zip(Qbc, golden_io.blocks_of(ent, "Qbc0")):
        np.testing.assert_array_equal(q, g)


def test_ramp_golden():
    _check_case(G5, "ramp")


def test_channel_two_steps_golden():
    npz = golden_io.load("g4_channel.npz")
    spec, arrays, ent = golden_io.case(npz, "channel")
    mesh, solver = cases.build(oracle, spec, arrays)
    Q = golden_io.blocks_of(ent, "Qin")
    t = 0.0
    for step in range(2):
        dt = float(solver.max_dt(Q))
        assert abs(dt - float(ent[f"dt{step}"])) <= 1e-13 * dt
        Q = solver.advance(Q, float(ent[f"dt{step}"]), t)
        t += dt
        for q, g in zip(Q, golden_io.blocks_of(ent, f"Q{step}")):
            assert golden_io.rel_err(q, g) <= 1e-13
    for (bname, face), (tau, qw) in solver.wall_stats(Q).items():
        np.testing.assert_allclose(tau, ent[f"tau_{bname}_{face}"], rtol=1e-11)
        np.testing.assert_allclose(qw, ent[f"qw_{bname}_{face}"], rtol=1e-10)

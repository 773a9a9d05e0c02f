"""Wall-load integration of the oracle (reference wmles/solver.py:181-232) against
closed forms, restating the reference's analytic check (tests/test_multiblock.py:161-190).

Uniform flow through a box of slip walls whose j faces are wall-modelled: every wall
face sees the same tau_w, wall pressure and tangent, so a face's force is
(p n + tau_w t) A and its moment about c is (x_f - c) x F, x_f the face's centroid.
The second case uses a box of unequal sides, a velocity with two tangential
components and an off-centre moment point, so that every force and moment component
is non-zero: a wrong sign, a swapped index of the cross product or a wrong area fails.
"""
import numpy as np

import oracle


def box_case(ns, L, n, uvw, T_wall=300.0):
    axes = [np.linspace(0.0, L[d], n[d] + 1) for d in range(3)]
    nodes = np.stack(np.meshgrid(*axes, indexing="ij"))
    blk = ns.Block("wall", nodes)
    mesh = ns.Mesh([blk])
    wall = ns.WallModel(ns.EquilibriumWallModel(), T_wall=T_wall)
    for face in ("imin", "imax", "kmin", "kmax"):
        blk.set_bc(face, ns.SlipWall())
    blk.set_bc("jmin", wall)
    blk.set_bc("jmax", wall)
    mesh.finalize()
    solver = ns.Solver(mesh, ns.Gas(), ns.KEP(2), sgs=None)

    def uniform(xc):
        z = 0.0 * xc[0]
        return np.stack([1.2 + z, uvw[0] + z, uvw[1] + z, uvw[2] + z, 1e5 + z])

    return solver, solver.init_state(uniform)


def check_reference_unit_box(ns, host=np.asarray):
    """tests/test_multiblock.py:161-190: unit cube, 4^3 cells, u = 30 m/s."""
    solver, Q = box_case(ns, (1.0, 1.0, 1.0), (4, 4, 4), (30.0, 0.0, 0.0))
    loads = solver.wall_loads(Q, (0.5, 0.5, 0.5))
    stats = solver.wall_stats(Q)
    assert set(loads) == {("wall", "jmin"), ("wall", "jmax")}
    for name, ny in (("jmin", -1.0), ("jmax", 1.0)):
        force, moment = (host(v) for v in loads[("wall", name)])
        np.testing.assert_allclose(force[0], np.sum(stats[("wall", name)][0]) / 16.0)
        np.testing.assert_allclose(force[1], ny * 1e5)
        np.testing.assert_allclose(force[2], 0.0, atol=1e-10)
        np.testing.assert_allclose(moment[:2], 0.0, atol=1e-8)
        np.testing.assert_allclose(moment[2], -ny * 0.5 * force[0])
    np.testing.assert_allclose(sum(host(v[1]) for v in loads.values()), 0.0, atol=1e-8)


def check_skew_box_closed_forms(ns, host=np.asarray):
    L, n, uvw = (1.0, 2.0, 3.0), (4, 5, 6), (30.0, 0.0, 10.0)
    c = np.array([0.2, -0.3, 0.7])
    solver, Q = box_case(ns, L, n, uvw)
    loads = solver.wall_loads(Q, tuple(c))
    stats = solver.wall_stats(Q)
    gas = oracle.Gas()
    # every face: h = half the wall-normal cell size, U = |u|, T_h = p / (rho R)
    h = L[1] / n[1] / 2.0
    U = float(np.linalg.norm(uvw))
    T = 1e5 / (1.2 * gas.R)
    tau, _ = oracle.EquilibriumWallModel()(gas, np.array([h]), np.array([U]), np.array([T]),
                                           np.array([1e5]), 300.0)
    t_hat = np.asarray(uvw) / U
    area = L[0] * L[2]
    for name, y_w, ny in (("jmin", 0.0, -1.0), ("jmax", L[1], 1.0)):
        face_tau = stats[("wall", name)][0]
        assert face_tau.shape == (n[0], n[2])
        np.testing.assert_allclose(face_tau, tau[0], rtol=1e-13)
        F = (1e5 * np.array([0.0, ny, 0.0]) + tau[0] * t_hat) * area
        M = np.cross(np.array([L[0] / 2, y_w, L[2] / 2]) - c, F)
        force, moment = (host(v) for v in loads[("wall", name)])
        np.testing.assert_allclose(force, F, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(moment, M, rtol=1e-12, atol=1e-9)
        assert np.all(np.abs(moment) > 1.0)          # every component exercised


def test_oracle_wall_loads_reference_unit_box():
    check_reference_unit_box(oracle)


def test_oracle_wall_loads_closed_forms_unequal_box():
    check_skew_box_closed_forms(oracle)


def test_oracle_wall_loads_rejects_a_bad_center():
    solver, Q = box_case(oracle, (1.0, 1.0, 1.0), (4, 4, 4), (30.0, 0.0, 0.0))
    for bad in ((0.0, 0.0), (0.0, np.nan, 0.0)):
        try:
            solver.wall_loads(Q, bad)
        except ValueError:
            continue
        raise AssertionError(f"moment centre {bad} accepted")

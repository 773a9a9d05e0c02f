"""Invariants and closed forms that pin the oracle without the reference
(reference tests `tests/test_multiblock.py:100-149`, `tests/test_sod.py`)."""
import numpy as np
import pytest

import cases
import oracle
from oracle.gas import Gas


def _build(kind, scheme, periodic_y=False, ybc=None, n=(12, 10, 8), sgs="vreman"):
    spec, arrays = cases.multiblock_spec(kind, n, scheme, ybc, periodic_y=periodic_y, sgs=sgs)
    return cases.build(oracle, spec, arrays)


@pytest.mark.parametrize("scheme", [{"kind": "KEP", "order": 2}, {"kind": "KEP", "order": 4},
                                    {"kind": "KEEP", "order": 4}, {"kind": "MUSCLHLLC"},
                                    {"kind": "Hybrid"}])
def test_freestream_preservation(scheme):
    """Uniform flow on a skewed periodic grid stays uniform (test_multiblock.py:100-107),
    here to a much tighter bound than the reference's 0.1."""
    mesh, s = _build("single", scheme, periodic_y=True)
    Q = s.init_state(lambda xc: np.stack([1.2 + 0 * xc[0], 50 + 0 * xc[0], -20 + 0 * xc[0],
                                          30 + 0 * xc[0], 1e5 + 0 * xc[0]]))
    R = s.rhs(Q)
    r = np.abs(mesh.blocks[0].interior(R[0])).max()
    assert r < 1e-6


def test_geometry_closure_and_volume():
    mesh, s = _build("single", {"kind": "KEP", "order": 4})
    g = mesh.geoms[0]
    Si, Sj, Sk = g.S
    close = (Si[:, 1:] - Si[:, :-1]) + (Sj[:, :, 1:] - Sj[:, :, :-1]) + (Sk[..., 1:] - Sk[..., :-1])
    assert np.abs(close).max() < 1e-15
    b = mesh.blocks[0]
    assert abs(b.interior(g.vol).sum() - np.prod(cases.L)) < 1e-12
    # exact volumes / centroids on an unskewed box
    x = np.linspace(0, 2, 5)
    nodes = np.stack(np.meshgrid(x, 0.5 * x, 0.25 * x, indexing="ij"))
    vol, xc, S, xf = oracle.compute_geometry(nodes)
    np.testing.assert_allclose(vol, 0.5 * 0.25 * 0.125, rtol=1e-14)
    np.testing.assert_allclose(S[0][0], 0.25 * 0.125, rtol=1e-14)
    np.testing.assert_allclose(xc[0][:, 0, 0], [0.25, 0.75, 1.25, 1.75], rtol=1e-14)


@pytest.mark.parametrize("kind", ["split", "flip", "rot_cyc"])
def test_multiblock_equals_single_block(kind):
    """RHS keyed by cell centre: re-oriented 2-block mesh == single block
    (test_multiblock.py:77-97)."""
    def gather(solver, R):
        out = {}
        for blk, r, g in zip(solver.blocks, R, solver.geoms):
            xc = blk.interior(g.xc).reshape(3, -1).T
            rr = blk.interior(r).reshape(5, -1).T
            for c, v in zip(xc, rr):
                out[tuple(np.round(c, 8))] = v
        return out
    m0, s0 = _build("single", {"kind": "Hybrid"})
    m1, s1 = _build(kind, {"kind": "Hybrid"})
    R0 = gather(s0, s0.rhs(s0.init_state(cases.prim_fn)))
    R1 = gather(s1, s1.rhs(s1.init_state(cases.prim_fn)))
    assert R0.keys() == R1.keys()
    scale = np.max(np.abs(np.array(list(R0.values()))))
    err = max(np.max(np.abs(R0[k] - R1[k])) for k in R0) / scale
    assert err < 1e-12


def test_characteristic_farfield_known_answers():
    """test_multiblock.py:110-131."""
    mesh, s = _build("single", {"kind": "KEP", "order": 4})
    blk, geom = mesh.blocks[0], mesh.geoms[0]
    ng, n = blk.ng, blk.n[1]
    prim_inf = np.array([1.2, 10.0, 0.0, 0.0, 1e5])
    bc = oracle.CharacteristicFarfield(prim_inf)

    def state(prim):
        return s.gas.prim_to_cons(np.asarray(prim, float).reshape(5, 1, 1, 1) * np.ones((1,) + blk.N))

    for q, side, expected in [(state([1.0, 0.0, 500.0, 0.0, 8e4]), 0, prim_inf),
                              (state([1.0, 0.0, 1000.0, 0.0, 8e4]), 1, [1.0, 0.0, 1000.0, 0.0, 8e4]),
                              (state(prim_inf), 0, prim_inf), (state(prim_inf), 1, prim_inf)]:
        filled = bc.fill(q, blk, geom, s.gas, 1, side)
        gj = ng - 1 if side == 0 else ng + n
        np.testing.assert_allclose(s.gas.cons_to_prim(filled[:, ng, gj, ng])[0], expected, rtol=1e-12)


def spalding_bisection(ut, nu, h, kap=0.41, B=5.2):
    yp = h * ut / nu
    lo, hi = 1e-6, 60.0
    f = lambda up: up + np.exp(-kap * B) * (np.exp(kap * up) - 1 - kap * up - (kap * up) ** 2 / 2
                                            - (kap * up) ** 3 / 6) - yp
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
    return ut * 0.5 * (lo + hi)


@pytest.mark.parametrize("ut,nu,h", [(2.0, 1.5e-5, 0.003), (0.5, 1.5e-5, 1e-4), (5.0, 1e-5, 0.01)])
def test_spalding_newton_matches_bisection(ut, nu, h):
    """test_multiblock.py:134-149."""
    U = spalding_bisection(ut, nu, h)
    rho = 1.2
    got = oracle.EquilibriumWallModel().utau(np.array([h]), np.array([U]), np.array([rho]),
                                             np.array([nu * rho]))
    assert abs(got[0] - ut) < 1e-6 * max(1.0, ut)


def test_kader_closed_form():
    gas = Gas()
    wm = oracle.EquilibriumWallModel()
    h, U, Th, ph, Tw = np.array([1e-3]), np.array([30.0]), np.array([310.0]), np.array([1e5]), 300.0
    tau, qw = wm(gas, h, U, Th, ph, Tw)
    rho_w = ph / (gas.R * Tw)
    mu_w = gas.mu(Tw)
    ut = np.sqrt(tau / rho_w)
    Pr = gas.Pr
    yp = h * ut * rho_w / mu_w
    G = 0.01 * (Pr * yp) ** 4 / (1 + 5 * Pr ** 3 * yp)
    beta = (3.85 * Pr ** (1 / 3) - 1.3) ** 2 + 2.12 * np.log(Pr)
    Tp = Pr * yp * np.exp(-G) + (2.12 * np.log(1 + yp) + beta) * np.exp(-1 / G)
    np.testing.assert_allclose(qw, rho_w * gas.cp * ut * (Tw - Th) / Tp, rtol=1e-12)
    assert qw[0] < 0          # wall colder than fluid: heat leaves the fluid
    tau_a, qw_a = wm(gas, h, U, Th, ph, None)
    assert qw_a[0] == 0.0


def test_sgs_and_ducros_special_cases():
    shape = (1, 1, 1)
    g_shear = np.zeros((3, 3) + shape)
    g_shear[0, 1] = 100.0                              # du/dy only: pure shear
    delta = np.full(shape, 0.01)
    rho = np.ones(shape)
    assert oracle.make_sgs("wale")(g_shear, delta, rho)[0, 0, 0] == 0.0
    assert oracle.make_sgs("vreman")(g_shear, delta, rho)[0, 0, 0] == 0.0
    smag = oracle.make_sgs("smagorinsky")(g_shear, delta, rho)[0, 0, 0]
    np.testing.assert_allclose(smag, (0.16 * 0.01) ** 2 * 100.0, rtol=1e-14)
    g_dil = np.zeros((3, 3) + shape)
    g_dil[0, 0] = g_dil[1, 1] = g_dil[2, 2] = 3.0
    np.testing.assert_allclose(oracle.ducros(g_dil), 1.0, rtol=1e-14)
    g_rot = np.zeros((3, 3) + shape)
    g_rot[0, 1], g_rot[1, 0] = 5.0, -5.0
    assert oracle.ducros(g_rot)[0, 0, 0] == 0.0


@pytest.mark.parametrize("fn", [oracle.kg_flux, oracle.numerics.keep2_flux])
def test_two_point_flux_of_constant_state_is_euler_flux(fn):
    w = np.array([1.1, 30.0, -4.0, 7.0, 9.5e4]).reshape(5, 1)
    S = np.array([0.3, -0.1, 0.2]).reshape(3, 1)
    g = 1.4
    F = fn(w, w, S, g)
    un = (w[1:4] * S).sum(0)
    E = w[4] / (g - 1) + 0.5 * w[0] * (w[1:4] ** 2).sum(0)
    ref = np.concatenate([w[0] * un, (w[0] * un * w[1:4] + w[4] * S)[:, 0], (E + w[4]) * un])
    np.testing.assert_allclose(F[:, 0], ref, rtol=1e-13)
    H = oracle.hllc(w, w, S, g)
    np.testing.assert_allclose(H[:, 0], ref, rtol=1e-12)


def test_viscous_flux_exact_for_linear_fields():
    """Cartesian grid, u = a y, T = T0 + b x: face stress mu a, heat flux -k b."""
    x = np.linspace(0, 1, 9)
    nodes = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    blk = oracle.Block("c", nodes)
    m = oracle.Mesh([blk])
    for f in ("imin", "imax", "jmin", "jmax", "kmin", "kmax"):
        blk.set_bc(f, oracle.SlipWall())
    m.finalize()
    geom = m.geoms[0]
    a, b = 40.0, 25.0
    xc = geom.xc
    phi = np.stack([a * xc[1], 0 * xc[0], 0 * xc[0], 300.0 + b * xc[0]])
    g = oracle.green_gauss(phi, geom)
    mu = np.full(blk.N, 2e-5)
    k = np.full(blk.N, 0.03)
    F = oracle.viscous_flux(blk, geom, 1, phi, g, mu, k)       # j faces, normal +y
    area = 1.0 / 64
    np.testing.assert_allclose(F[1], -2e-5 * a * area, rtol=1e-10)
    yf = 0.5 * (blk.slab(xc[1], 1, -1, 0) + blk.slab(xc[1], 1, 0, 1))
    np.testing.assert_allclose(F[4], -(a * yf) * 2e-5 * a * area, rtol=1e-9, atol=1e-14)
    F0 = oracle.viscous_flux(blk, geom, 0, phi, g, mu, k)      # i faces, normal +x
    np.testing.assert_allclose(F0[2], -2e-5 * a * area, rtol=1e-10)
    np.testing.assert_allclose(F0[4], -0.03 * b * area, rtol=1e-10)
    assert np.abs(F0[1]).max() < 1e-15 and np.abs(F0[0]).max() == 0.0


def _sod(scheme, nx=100, tend=0.2):
    """test_sod.py:7-28 setup with the oracle."""
    gas = Gas(R=1.0, mu_ref=1e-9)
    x = np.linspace(0, 1, nx + 1)
    y = np.linspace(0, 0.1, 4)
    b = oracle.Block("tube", np.stack(np.meshgrid(x, y, y, indexing="ij")))
    m = oracle.Mesh([b])
    b.set_bc("imin", oracle.Extrapolate())
    b.set_bc("imax", oracle.Extrapolate())
    m.connect(b, "jmin", b, "jmax", shift=(0, -0.1, 0))
    m.connect(b, "kmin", b, "kmax", shift=(0, 0, -0.1))
    m.finalize()
    s = oracle.Solver(m, gas, scheme, sgs=None, cfl=0.5)
    Q = s.init_state(lambda xc: np.stack([np.where(xc[0] < 0.5, 1.0, 0.125), 0 * xc[0], 0 * xc[0],
                                          0 * xc[0], np.where(xc[0] < 0.5, 1.0, 0.1)]))
    mass0 = float(np.sum(b.interior(Q[0][0]) * b.interior(m.geoms[0].vol)))
    t = 0.0
    while t < tend - 1e-12:
        dt = min(float(s.max_dt(Q)), tend - t)
        Q = s.advance(Q, dt)
        t += dt
    w, _ = gas.cons_to_prim(Q[0])
    mass1 = float(np.sum(b.interior(Q[0][0]) * b.interior(m.geoms[0].vol)))
    return b.interior(m.geoms[0].xc)[0, :, 1, 1], b.interior(w)[:, :, 1, 1], mass0, mass1


@pytest.mark.parametrize("scheme", [oracle.MUSCLHLLC(), oracle.Hybrid(th0=0.3, th1=0.6)])
def test_sod_plateaus_match_exact_riemann(scheme):
    """Exact solution at t=0.2: p*=0.30313, u*=0.92745, rho*L=0.42632, rho*R=0.26557."""
    xc, w, m0, m1 = _sod(scheme)
    i = np.argmin(abs(xc - 0.75))
    j = np.argmin(abs(xc - 0.60))
    assert abs(w[4, i] - 0.30313) < 0.01
    assert abs(w[1, i] - 0.92745) < 0.03
    assert abs(w[0, i] - 0.26557) < 0.01
    assert abs(w[0, j] - 0.42632) < 0.02
    assert abs(m1 - m0) < 1e-12         # mass conserved (extrapolated ends carry no net flux yet)

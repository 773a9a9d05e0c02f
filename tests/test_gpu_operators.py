"""The reference's operator callables on the device against the oracle's restatement of
the same callables, on ghost-filled multiblock states with interior noise:

* scheme(blk, geom, w, gas, d, theta) -> F for KEP(2/4), KEEP(4), MUSCLHLLC (albada,
  minmod) and Hybrid (reference wmles/schemes.py:3-5,57-166);
* make_sgs(name)(g, delta, rho) -> mu_t for None, Smagorinsky, WALE and Vreman
  (reference wmles/sgs.py:45-50);
* WallModel.wall_flux(w, T, blk, geom, gas, axis, side) -> (F, tau_w, q_w), isothermal
  and adiabatic, n_match 1 and 2 (reference wmles/bc.py:145-166); also through the
  callable-model path with a user wall model.

Bar: max|diff| / max|oracle| per component <= 1e-12 (the FP64 form's RHS bar)."""
import numpy as np
import pytest

import cases
import oracle
from oracle import numerics as ON

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402

RTOL = 1e-12


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.ndim == 0 or b.shape[0] != 5:
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return max(float(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1e-300)) for c in range(5))


def _state(kind, ybc=None, seed=3):
    spec, arrays = cases.multiblock_spec(kind, (12, 10, 8), None, ybc)
    m0, s0 = cases.build(oracle, spec, arrays)
    m1, s1 = cases.build(wmx, spec, arrays)
    Q = cases.interior_noise(m0.blocks, s0.init_state(cases.prim_fn), seed, 0.05)
    Qbc = s0.apply_bcs(Q, m0.geoms)
    return m0, s0, m1, s1, Qbc


@pytest.mark.parametrize("kind", ["single", "rot_cyc"])
def test_scheme_callables_match_the_oracle(kind):
    m0, s0, m1, s1, Qbc = _state(kind)
    gas0, gas1 = s0.gas, s1.gas
    pairs = [(ON.KEP(2), wmx.KEP(2)), (ON.KEP(4), wmx.KEP(4)), (ON.KEEP(4), wmx.KEEP(4)),
             (ON.MUSCLHLLC(), wmx.MUSCLHLLC()), (ON.MUSCLHLLC("minmod"), wmx.MUSCLHLLC("minmod")),
             (ON.Hybrid(), wmx.Hybrid()), (ON.Hybrid(ON.KEEP(2), None, 0.2, 0.5),
                                           wmx.Hybrid(wmx.KEEP(2), None, 0.2, 0.5))]
    for blk0, blk1, g0, g1, q in zip(m0.blocks, m1.blocks, m0.geoms, m1.geoms, Qbc):
        w, T = gas0.cons_to_prim(q)
        phi = np.concatenate([w[1:4], T[None]], axis=0)
        theta = ON.ducros(ON.green_gauss(phi, g0)[:3])
        print(kind, blk0.name, "theta range", float(theta.min()), float(theta.max()), flush=True)
        for sc0, sc1 in pairs:
            for d in range(3):
                F0 = sc0(blk0, g0, w, gas0, d, theta)
                F1 = sc1(blk1, g1, w, gas1, d, theta)
                err = _rel(F1, F0)
                assert err <= RTOL, (kind, type(sc0).__name__, d, err)


def test_sgs_callables_match_the_oracle():
    m0, s0, m1, s1, Qbc = _state("split")
    for blk0, g0, q in zip(m0.blocks, m0.geoms, Qbc):
        w, T = s0.gas.cons_to_prim(q)
        phi = np.concatenate([w[1:4], T[None]], axis=0)
        g = ON.green_gauss(phi, g0)[:3]
        delta = g0.vol ** (1.0 / 3.0)
        for name, kw in ((None, {}), ("smagorinsky", {}), ("wale", {"Cw": 0.5}), ("vreman", {})):
            ref = ON.make_sgs(name, **kw)(g, delta, w[0])
            got = wmx.make_sgs(name, **kw)(g, delta, w[0])
            err = _rel(got, np.broadcast_to(ref, w[0].shape))
            assert err <= RTOL, (name, err)


@pytest.mark.parametrize("T_wall,n_match", [(300.0, 1), (None, 2)])
def test_wall_flux_matches_the_oracle(T_wall, n_match):
    ybc = {"kind": "WallModel", "T_wall": T_wall, "n_match": n_match}
    m0, s0, m1, s1, Qbc = _state("flip", ybc)
    done = 0
    for blk0, blk1, g0, g1, q in zip(m0.blocks, m1.blocks, m0.geoms, m1.geoms, Qbc):
        w, T = s0.gas.cons_to_prim(q)
        for (axis, side), bc0 in blk0.faces.items():
            if not isinstance(bc0, oracle.WallModel):
                continue
            bc1 = blk1.faces[(axis, side)]
            F0, tau0, qw0 = bc0.wall_flux(w, T, blk0, g0, s0.gas, axis, side)
            F1, tau1, qw1 = bc1.wall_flux(w, T, blk1, g1, s1.gas, axis, side)
            assert _rel(F1, F0) <= RTOL and _rel(tau1, tau0) <= RTOL
            if T_wall is not None:
                assert _rel(qw1, qw0) <= RTOL
            else:
                assert float(torch.abs(qw1).max()) == 0.0 and np.abs(qw0).max() == 0.0
            # the callable-model path (any model(gas, h, U, T_h, p_h, T_wall)) with the
            # equilibrium model wrapped as a user callable gives the same result
            user = wmx.WallModel(lambda gas, h, U, Th, ph, Tw: bc1.model(gas, h, U, Th, ph, Tw),
                                 T_wall=bc1.T_wall, n_match=bc1.n_match)
            F2, tau2, qw2 = user.wall_flux(w, T, blk1, g1, s1.gas, axis, side)
            assert _rel(F2, F0) <= RTOL and _rel(tau2, tau0) <= RTOL
            done += 1
    assert done == 4

"""User Python boundary conditions and sources (the reference's BC protocol
`fill(q, blk, geom, gas, axis, side)`, wmles/bc.py:1-8,28-35, and source protocol
`src(blk, q, geom, gas) -> (5, n_i, n_j, n_k)`, wmles/solver.py:103-104) on the device:
the solver runs the HIP sweep phase by phase with the user fills at their places in the
reference's order, the HIP RHS of the filled state, the user sources, and the SSP-RK3
stage combinations in torch.  Compared with the oracle solver holding the same BC and
source: ghosts bit for bit (the user fill here is the oracle's slip-wall fill), RHS and
advance within 1e-12 (max-normalised per component)."""
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


def host(Q):
    return [q.detach().cpu().numpy() for q in Q]


class HostSlip(wmx.BC):
    """A user BC: the oracle's slip-wall fill on a host copy of the block's state."""

    def __init__(self, lookup):
        self.lookup = lookup                     # block name -> (oracle block, oracle geom)
        self.calls = 0

    def fill(self, q, blk, geom, gas, axis, side):
        ob, og = self.lookup[blk.name]
        self.calls += 1
        out = oracle.SlipWall().fill(q.detach().cpu().numpy(), ob, og, oracle.Gas(), axis, side)
        return torch.as_tensor(np.asarray(out), device=q.device)


def _body_force(lib):
    """src(blk, q, geom, gas): a body force along x proportional to rho, with its work."""
    def src(blk, q, geom, gas):
        rho = blk.interior(q[0])
        ru = blk.interior(q[1])
        f = 250.0 * rho
        z = 0.0 * rho
        return lib.stack([z, f, z, z, 250.0 * ru])
    return src


@pytest.mark.parametrize("kind", ["split", "rot_cyc"])
def test_user_bc_and_source_match_the_oracle(kind):
    spec, arrays = cases.multiblock_spec(kind, (12, 10, 8), {"kind": "KEP", "order": 4},
                                         {"kind": "SlipWall"},
                                         constraints=[{"kind": "ConstantMassFlux", "direction": 0,
                                                       "target": 30.0}])
    m0, s0 = cases.build(oracle, spec, arrays, solver_kwargs={"sources": [_body_force(np)]})
    lookup = {b.name: (b, g) for b, g in zip(m0.blocks, m0.geoms)}
    user = HostSlip(lookup)
    spec1 = dict(spec)
    spec1["bcs"] = [[nm, face, {"kind": "Object", "obj": user}] for nm, face, _ in spec["bcs"]]
    m1, s1 = cases.build(wmx, spec1, arrays, solver_kwargs={"sources": [_body_force(torch)]})
    assert s1._external and s1._fallback
    Q = cases.interior_noise(m0.blocks, s0.init_state(cases.prim_fn), 5, 0.05)
    Q = cases.noisy_ghost_state(m0.blocks, Q, 9)
    # the sweep: ghosts bit for bit
    for a, b in zip(host(s1.apply_bcs(Q, None, 0.0)), s0.apply_bcs(Q, m0.geoms, 0.0)):
        assert np.array_equal(a, b)
    assert user.calls == 2 * 2 * len(m1.blocks)          # bc_passes x 2 faces per block
    R0, R1 = s0.rhs(Q), host(s1.rhs(Q))
    for a, b in zip(R1, R0):
        assert golden_io.rel_err(a, b) <= 1e-12
    dt = float(s0.max_dt(Q))
    assert float(s1.max_dt(Q)) == dt
    A0, A1 = s0.advance(Q, dt), host(s1.advance(Q, dt))
    for a, b in zip(A1, A0):
        assert golden_io.rel_err(a, b) <= 1e-12
    # step = max_dt + advance on the torch path as well
    S1, dts = s1.step(Q)
    assert float(dts) == dt
    for a, b in zip(host(S1), A1):
        assert np.array_equal(a, b)
    # the inputs are unchanged
    Qd = tuple(torch.as_tensor(q, device="cuda") for q in Q)
    s1.advance(Qd, dt)
    for a, b in zip(host(Qd), Q):
        assert np.array_equal(a, b)

"""Allocation-free CFL time step: wmx_max_dt_noalloc (static device scratch for the
per-workgroup partials) called through the C ABI must return wmx_max_dt's dt bit for
bit, on two calls in a row, the second on a smaller case, and also when a plan has no
workspace bound."""
import ctypes as C

import pytest

import cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402
from wmles_jax_amd import _lib  # noqa: E402


def _case(nblocks, nx, ny, nz, Re_b, seed, precision="fp64"):
    spec, arrays = cases.channel_spec(nblocks, nx, ny, nz, {"kind": "KEP", "order": 2}, "wale",
                                      Re_b=Re_b)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": precision})
    pr = spec["props"]
    Q = s.apply_bcs(s.init_state(lambda xc: cases.channel_prim(xc, pr, seed)))
    return s, Q


def _dt(fn, plan, Q):
    out = torch.full((16,), float("nan"), dtype=torch.float64, device=Q[0].device)
    _lib.check(fn(plan, _lib.ptr_array(Q), C.c_void_p(out.data_ptr()), _lib.stream_ptr()))
    torch.cuda.synchronize()
    return out


def test_noalloc_max_dt_equals_max_dt_on_two_calls():
    big, Qb = _case(2, 96, 48, 32, 2e4, 0)          # 147,456 interior cells: 144 partials
    small, Qs = _case(1, 24, 16, 8, 8e3, 1)         # 3,072 interior cells: 3 partials
    lib = _lib.lib
    for s, Q in ((big, Qb), (small, Qs)):           # the second call on the smaller case
        ref = _dt(lib.wmx_max_dt, s._plan, Q)
        got = _dt(lib.wmx_max_dt_noalloc, s._plan, Q)
        assert float(ref[0]) > 0.0
        assert torch.equal(got[:1], ref[:1]), (float(got[0]), float(ref[0]))
    # the state and the first plan again: the scratch keeps no state across calls
    assert torch.equal(_dt(lib.wmx_max_dt_noalloc, big._plan, Qb)[:1], _dt(lib.wmx_max_dt, big._plan, Qb)[:1])


def test_noalloc_max_dt_without_workspace_and_per_case():
    """A plan with no workspace bound (created directly through the C ABI) gives the same
    dt; an ensemble plan gives one dt per case."""
    s, Q = _case(2, 40, 24, 16, 2e4, 3)
    plan = C.c_void_p()
    inflows = (_lib.InflowDesc * 1)()
    _lib.check(_lib.lib.wmx_plan_create(len(s._descs), s._descs, C.byref(s._nd), 0, inflows,
                                         C.byref(plan)))
    try:
        got = _dt(_lib.lib.wmx_max_dt_noalloc, plan, Q)
        ref = _dt(_lib.lib.wmx_max_dt, s._plan, Q)
        assert torch.equal(got[:1], ref[:1])
    finally:
        _lib.lib.wmx_plan_destroy(plan)
    a, Qa = _case(2, 32, 16, 16, 2e4, 0)
    b, Qbb = _case(3, 48, 16, 8, 8e3, 1)
    ens = wmx.EnsembleSolver([a, b])
    Qe = tuple(Qa) + tuple(Qbb)
    ref = _dt(_lib.lib.wmx_max_dt, ens._plan, Qe)
    got = _dt(_lib.lib.wmx_max_dt_noalloc, ens._plan, Qe)
    assert torch.equal(got[:2], ref[:2])
    assert float(got[0]) == float(_dt(_lib.lib.wmx_max_dt, a._plan, Qa)[0])
    assert float(got[1]) == float(_dt(_lib.lib.wmx_max_dt, b._plan, Qbb)[0])

"""Persistent stages (one cooperative launch per RK stage: sweep phases, primitives,
gradients with the dt min and wall faces, flux tiles, behind grid-wide barriers) against
the separate launches: the same dt and the same states, bit for bit, for the 16-case
uneven FP32-bulk ensemble (KEP2 + WALE) and for single cases of the other compiled scheme
configurations (Hybrid KEP4 + MUSCL-HLLC with Vreman, KEP4 with Smagorinsky, KEEP with
no SGS), through step (the stage-1 CFL dt) and through advance with a host dt; plans the
persistent form does not cover are refused."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402


def _solvers(specs, precision="fp32"):
    solvers, fns = [], []
    for spec, arrays, pr, seed in specs:
        _, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": precision})
        solvers.append(s)
        fns.append(lambda xc, pr=pr, seed=seed: cases.channel_prim(xc, pr, seed))
    return solvers, fns


def test_persistent_ensemble16_equals_separate_launches():
    specs = cases.ensemble_specs(16)
    print("[persistent] building the 16-case ensemble", flush=True)
    solvers, fns = _solvers(specs)
    eager, pers = wmx.EnsembleSolver(solvers), wmx.EnsembleSolver(solvers)
    pers.set_persistent(True)
    Q0 = eager.init_state(fns)
    A, B = Q0, Q0
    for it in range(3):
        A, dta = eager.step(A)
        B, dtb = pers.step(B, graph=True)           # (graphs are not used while persistent)
        torch.cuda.synchronize()
        assert torch.equal(dta, dtb), it
        for c in range(16):
            for a, b in zip(A[c], B[c]):
                assert torch.equal(a, b), (it, c)
        print(f"[persistent] ensemble16 step {it} equal", flush=True)


@pytest.mark.parametrize("scheme,sgs", [
    ({"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}}, "vreman"),
    ({"kind": "KEP", "order": 4}, "smagorinsky"),
    ({"kind": "KEEP", "order": 2}, None)])
def test_persistent_single_case_equals_separate_launches(scheme, sgs):
    spec, arrays = cases.channel_spec(2, 48, 24, 16, scheme, sgs, Re_b=2e4, splits=[0, 20, 48])
    _, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    _, sp = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    sp.set_persistent(True)
    pr = spec["props"]
    Q0 = s.init_state(lambda xc: cases.channel_prim(xc, pr, 4))
    A, B = Q0, Q0
    for it in range(3):
        A, dta = s.step(A)
        B, dtb = sp.step(B)
        assert float(dta) == float(dtb)
        for a, b in zip(A, B):
            assert torch.equal(a, b), it
    dt = 0.7 * float(dta)                           # advance with a host dt (no CFL partials)
    for a, b in zip(s.advance(A, dt), sp.advance(B, dt)):
        assert torch.equal(a, b)
    sp.set_persistent(False)                        # back to the separate launches
    for a, b in zip(s.advance(A, dt), sp.advance(B, dt)):
        assert torch.equal(a, b)


def test_persistent_refuses_uncovered_plans():
    spec, arrays = cases.channel_spec(2, 32, 16, 16, {"kind": "KEP", "order": 2}, "wale")
    _, s64 = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp64"})
    with pytest.raises(RuntimeError, match="FP32-bulk"):
        s64.set_persistent(True)
    rspec, rarrays, _ = cases.ramp_spec(wmx, 24, 16, 8, n_modes=4)
    _, sr = cases.build(wmx, rspec, rarrays, solver_kwargs={"precision": "fp32"})
    with pytest.raises(RuntimeError, match="inflow"):
        sr.set_persistent(True)

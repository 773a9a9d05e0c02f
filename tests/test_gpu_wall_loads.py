"""Solver.wall_loads on the device (tau_w from the HIP wall kernel, integration with
device tensors) against the closed forms the oracle is pinned by, and against the
oracle on a noisy two-block wall-modelled channel (reference wmles/solver.py:181-232)."""
import numpy as np
import pytest

import cases
import oracle
from test_oracle_wall_loads import check_reference_unit_box, check_skew_box_closed_forms

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a HIP device", allow_module_level=True)

import wmles_jax_amd as wmx  # noqa: E402


def test_wall_loads_reference_unit_box():
    check_reference_unit_box(wmx)


def test_wall_loads_closed_forms_unequal_box():
    check_skew_box_closed_forms(wmx)


def test_wall_loads_match_the_oracle_on_a_channel():
    spec, arrays = cases.channel_spec(2, 24, 16, 12, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", splits=[0, 10, 24], n_match=2)
    m0, s0 = cases.build(oracle, spec, arrays)
    m1, s1 = cases.build(wmx, spec, arrays)
    pr = spec["props"]
    Q = cases.interior_noise(m0.blocks, s0.init_state(lambda xc: cases.channel_prim(xc, pr, 3)), 4, 0.05)
    center = (0.3 * pr["d"], 0.9 * pr["d"], -0.2 * pr["d"])
    ref = s0.wall_loads(Q, center)
    got = s1.wall_loads(Q, center)
    assert set(ref) == set(got) and len(ref) == 4
    for key in ref:
        for a, b in zip(got[key], ref[key]):
            err = np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b)
            print(key, err)
            assert err <= 1e-12, (key, err)
    with pytest.raises(ValueError):
        s1.wall_loads(Q, (0.0, 0.0))

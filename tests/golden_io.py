"""Load the golden vectors written by tests/golden/make_golden.py."""
import os

import numpy as np

import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(fname):
    return np.load(os.path.join(GOLDEN, fname), allow_pickle=False)


def case_keys(npz):
    return sorted({k.split("__")[0] for k in npz.files})


def case(npz, key):
    """-> (spec, arrays, entries) where entries maps the remaining names to arrays."""
    pre = key + "__"
    spec = cases.loads(npz[pre + "spec"])
    arrays, entries = {}, {}
    for k in npz.files:
        if not k.startswith(pre) or k == pre + "spec":
            continue
        name = k[len(pre):]
        if name.startswith("nodes_") or name == "inflow_profile":
            arrays[name] = npz[k]
        else:
            entries[name] = npz[k]
    return spec, arrays, entries


def blocks_of(entries, prefix):
    out, i = [], 0
    while f"{prefix}_{i}" in entries:
        out.append(entries[f"{prefix}_{i}"])
        i += 1
    return tuple(out)


def rel_err(a, b):
    """max over components of max|a - b| / max|b| (component-normalised, as
    tests/test_multiblock.py:94-95 normalises by max|R|)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    worst = 0.0
    for c in range(a.shape[0]):
        scale = np.max(np.abs(b[c]))
        if scale == 0.0:
            scale = 1.0
        worst = max(worst, float(np.max(np.abs(a[c] - b[c])) / scale))
    return worst

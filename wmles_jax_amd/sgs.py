"""Eddy-viscosity SGS models (reference `wmles/sgs.py`): descriptors evaluated by the
HIP gradient kernel."""
from . import _lib

_DEFAULT = {"smagorinsky": ("Cs", 0.16, _lib.SGS_SMAGORINSKY),
            "wale": ("Cw", 0.325, _lib.SGS_WALE),
            "vreman": ("Cs", 0.17, _lib.SGS_VREMAN)}


class SGSModel:
    def __init__(self, name, coef):
        self.name, self.coef = name, coef
        self.code = _DEFAULT[name][2] if name else _lib.SGS_NONE

    def __call__(self, g, delta, rho):
        raise NotImplementedError("mu_t is evaluated inside the HIP step")


def make_sgs(name="vreman", **kw):
    """name in {None, 'none', 'smagorinsky', 'wale', 'vreman'} (sgs.py:45-50)."""
    if name in (None, "none"):
        return SGSModel(None, 0.0)
    key, default, _ = _DEFAULT[name]
    unknown = set(kw) - {key}
    if unknown:
        raise TypeError(f"{name}() got unexpected arguments {sorted(unknown)}")
    return SGSModel(name, float(kw.get(key, default)))

"""Eddy-viscosity SGS models (reference `wmles/sgs.py`): descriptors the HIP gradient
kernel evaluates inside the step, and the reference's callable `f(g, delta, rho) -> mu_t`
(`wmles/sgs.py:45-50`) through `wmx_sgs_eval` (the same FP64 device code)."""
import ctypes as C

import torch

from . import _lib
from .backend import asarray

_DEFAULT = {"smagorinsky": ("Cs", 0.16, _lib.SGS_SMAGORINSKY),
            "wale": ("Cw", 0.325, _lib.SGS_WALE),
            "vreman": ("Cs", 0.17, _lib.SGS_VREMAN)}


class SGSModel:
    def __init__(self, name, coef):
        self.name, self.coef = name, coef
        self.code = _DEFAULT[name][2] if name else _lib.SGS_NONE

    def __call__(self, g, delta, rho):
        """mu_t = rho nu_t(g, delta): g (3, 3, ...) with g[i, j] = du_i/dx_j; delta and rho
        broadcast to g's trailing shape."""
        from .gas import Gas
        from .solver import numerics_desc
        g = asarray(g)
        if g.ndim < 2 or tuple(g.shape[:2]) != (3, 3):
            raise ValueError(f"g must be (3, 3, ...), got {tuple(g.shape)}")
        shape = tuple(g.shape[2:])
        delta = torch.broadcast_to(asarray(delta), shape).contiguous()
        rho = torch.broadcast_to(asarray(rho), shape).contiguous()
        g = g.reshape(9, -1).contiguous()
        out = torch.empty(shape, dtype=torch.float64, device=g.device)
        nd = numerics_desc(Gas(), None, self, 0.8, 2, (), ())
        _lib.check(_lib.lib.wmx_sgs_eval(C.byref(nd), int(g.shape[1]), C.c_void_p(g.data_ptr()),
                                         C.c_void_p(delta.data_ptr()), C.c_void_p(rho.data_ptr()),
                                         C.c_void_p(out.data_ptr()), _lib.stream_ptr()))
        return out


def make_sgs(name="vreman", **kw):
    """name in {None, 'none', 'smagorinsky', 'wale', 'vreman'} (sgs.py:45-50)."""
    if name in (None, "none"):
        return SGSModel(None, 0.0)
    key, default, _ = _DEFAULT[name]
    unknown = set(kw) - {key}
    if unknown:
        raise TypeError(f"{name}() got unexpected arguments {sorted(unknown)}")
    return SGSModel(name, float(kw.get(key, default)))

"""Device selection and the host array module user callbacks may use.

The reference switches between JAX and NumPy (`wmles/backend.py`); here the state
lives in torch tensors on the HIP device and every step kernel is hand-written HIP.
`xp` is NumPy so that initial-condition callbacks written for the reference
(`prim_fn(xc)` with `xp.sin`, ...) run unchanged on host copies of the cell centres.
"""
import os

import numpy as xp  # noqa: F401  (host arrays for user callbacks)
import torch

HAS_JAX = False
_REQ = os.environ.get("WMLES_DEVICE", "auto").strip().lower()


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("wmles_jax_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def device_info():
    d = device()
    return f"hip:{torch.cuda.get_device_name(d)}:{d}"


def asarray(value, dtype=torch.float64):
    """Host or device array -> contiguous float64 tensor on the HIP device."""
    if isinstance(value, torch.Tensor):
        return value.to(device=device(), dtype=dtype).contiguous()
    return torch.as_tensor(xp.ascontiguousarray(value), dtype=dtype, device=device())

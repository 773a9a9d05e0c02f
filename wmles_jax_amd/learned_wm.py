"""Learned equilibrium wall model (no reference code).

It plugs in at the reference's wall-model callable boundary
`model(gas, h, U, T_h, p_h, T_wall) -> (tau_w, q_w)` (`wmles/wall_model.py:3-9`, called
at `wmles/bc.py:161`): pass it as `WallModel(LearnedWallModel(...), ...)` and the
solver's wall-flux kernel evaluates it in place of the Newton solve.

Network 3 -> 64 -> 64 -> 2, hidden activation a / sqrt(1 + a^2) (ISRU: one hardware
reciprocal square root per unit).  Inputs: log Re_h, h/Delta
(Delta = vol^(1/3) of the sampled cell), T_h/T_w; outputs log u+ (u+ = U/u_tau) and
log T+ (Kader's T+ at y+ = Re_h/u+), both normalised.  Everything runs in HIP kernels
of libwmles_amd.so:

* training data: the method itself (`wmx_wm_dataset`: the reference's 30-step
  Newton for Spalding's law, Kader's closed form), features uniform over the trained
  box, deterministic in (seed, index);
* training: FP32 forward + backward kernels and Adam (`wmx_mlp_train_step`);
* inference: BF16 weights and activations on v_mfma_f32_16x16x16bf16_1k with FP32
  accumulation (`wmx_wall_net_eval`, and the plan's wall kernel after
  `wmx_plan_set_wall_net`).  The network's u+ starts `n_polish` FP64 Newton steps; a
  point falls back to the reference's full Newton solve when log Re_h or T_h/T_w lies
  outside the trained range or the relative Spalding residual exceeds `tol`.
"""
import ctypes as C
import math
import time

import numpy as np
import torch

from . import _lib
from .wall_model import EquilibriumWallModel

_NPARAM, _NBF16, _NF32 = (C.c_int32(), C.c_int32(), C.c_int32())
_lib.check(_lib.lib.wmx_mlp_sizes(C.byref(_NPARAM), C.byref(_NBF16), C.byref(_NF32)))
NPARAM, NBF16, NF32 = _NPARAM.value, _NBF16.value, _NF32.value
H, NIN, NOUT = 64, 3, 2
_O_W1, _O_W2, _O_W3 = 0, H * NIN + H, H * NIN + H + H * H + H


def _nd(gas, base):
    from .solver import numerics_desc
    nd = numerics_desc(gas, None, None, 0.8, 2, (), ())
    nd.kappa, nd.B, nd.n_iter = base.kappa, base.B, base.n_iter
    nd.wm_ekb, nd.wm_pr13, nd.wm_pr3, nd.wm_beta = base.host_constants(gas)
    return nd


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class LearnedWallModel:
    """Learned model with the Newton model `base` as its data source and fallback.

    re_range / hd_range / tr_range: trained box of Re_h, h/Delta, T_h/T_w.  Points
    outside the Re_h or T_h/T_w range take the Newton fallback; h/Delta is clamped into
    its range (the equilibrium laws do not depend on it).
    n_polish: Newton steps from the network's u+.  tol: fallback when
    |F(u+)|/u+ > tol after the polish."""

    def __init__(self, base=None, re_range=(1e2, 1e6), hd_range=(0.5, 4.0), tr_range=(0.8, 1.25),
                 n_polish=3, tol=1e-6, device=None):
        self.base = base or EquilibriumWallModel()
        self.kappa, self.B, self.n_iter = self.base.kappa, self.base.B, self.base.n_iter
        self.lo = (math.log(re_range[0]), float(hd_range[0]), float(tr_range[0]))
        self.hi = (math.log(re_range[1]), float(hd_range[1]), float(tr_range[1]))
        self.x_center = tuple(0.5 * (a + b) for a, b in zip(self.lo, self.hi))
        self.x_scale = tuple(0.5 * (b - a) for a, b in zip(self.lo, self.hi))
        self.y_center, self.y_scale = (0.0, 0.0), (1.0, 1.0)
        self.n_polish, self.tol = int(n_polish), float(tol)
        self.device = torch.device(device or "cuda")
        self.params = None
        self.w_bf16 = torch.zeros(NBF16, dtype=torch.int16, device=self.device)
        self.w_fp32 = torch.zeros(NF32, dtype=torch.float32, device=self.device)
        self.fallbacks = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.trained = False
        self.history = {}

    # ------------------------------------------------------------------ descriptor
    def desc(self, count_fallbacks=True):
        d = _lib.WallNetDesc()
        d.w_bf16 = self.w_bf16.data_ptr()
        d.w_fp32 = self.w_fp32.data_ptr()
        for k in range(3):
            d.x_center[k], d.x_scale[k] = self.x_center[k], self.x_scale[k]
            d.lo[k], d.hi[k] = self.lo[k], self.hi[k]
        for k in range(2):
            d.y_center[k], d.y_scale[k] = self.y_center[k], self.y_scale[k]
        d.n_polish, d.tol = self.n_polish, self.tol
        d.fallback_count = self.fallbacks.data_ptr() if count_fallbacks else None
        return d

    # ----------------------------------------------------------------- parameters
    def init_params(self, seed=0):
        """Xavier-uniform weights, zero biases."""
        rng = np.random.default_rng(seed)
        P = np.zeros(NPARAM, dtype=np.float32)
        for off, fo, fi in ((_O_W1, H, NIN), (_O_W2, H, H), (_O_W3, NOUT, H)):
            a = math.sqrt(6.0 / (fi + fo))
            P[off:off + fo * fi] = rng.uniform(-a, a, fo * fi).astype(np.float32)
        self.set_params(P)

    def set_params(self, P):
        self.params = torch.as_tensor(np.asarray(P, dtype=np.float32)).to(self.device).contiguous()
        self.pack()

    def pack(self, stream=None):
        _lib.check(_lib.lib.wmx_mlp_pack(_ptr(self.params), _ptr(self.w_bf16), _ptr(self.w_fp32),
                                         _lib.stream_ptr(stream)))

    def state_dict(self):
        return {"params": self.params.cpu().numpy(), "lo": np.array(self.lo), "hi": np.array(self.hi),
                "y_center": np.array(self.y_center), "y_scale": np.array(self.y_scale),
                "n_polish": self.n_polish, "tol": self.tol}

    def load_state_dict(self, sd):
        self.lo, self.hi = tuple(map(float, sd["lo"])), tuple(map(float, sd["hi"]))
        self.x_center = tuple(0.5 * (a + b) for a, b in zip(self.lo, self.hi))
        self.x_scale = tuple(0.5 * (b - a) for a, b in zip(self.lo, self.hi))
        self.y_center = tuple(map(float, sd["y_center"]))
        self.y_scale = tuple(map(float, sd["y_scale"]))
        self.n_polish, self.tol = int(sd["n_polish"]), float(sd["tol"])
        self.set_params(sd["params"])
        self.trained = True

    def save(self, path):
        np.savez(path, **self.state_dict())

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            self.load_state_dict({k: f[k] for k in f.files})

    # --------------------------------------------------------------------- data
    def dataset(self, gas, n, seed=0, raw=False, stream=None):
        """n samples of the method: X (n,3), Y (n,2) normalised FP32; raw (n,5) FP64
        (log Re_h, h/Delta, T_h/T_w, u+, T+) when asked."""
        X = torch.empty((n, 3), dtype=torch.float32, device=self.device)
        Y = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        R = torch.empty((n, 5), dtype=torch.float64, device=self.device) if raw else None
        nd = _nd(gas, self.base)
        _lib.check(_lib.lib.wmx_wm_dataset(C.byref(nd), C.byref(self.desc(False)), n, seed, _ptr(X),
                                           _ptr(Y), _ptr(R), _lib.stream_ptr(stream)))
        return (X, Y, R) if raw else (X, Y)

    # ------------------------------------------------------------------ training
    def grad(self, X, Y, offset=0, batch=None, stream=None):
        """(summed-group gradient, loss) of the mean squared error over samples
        [offset, offset + batch); parameters unchanged."""
        batch = X.shape[0] - offset if batch is None else batch
        ng = (batch + 63) // 64
        part = torch.empty(ng * NPARAM, dtype=torch.float32, device=self.device)
        lpart = torch.empty(ng, dtype=torch.float32, device=self.device)
        g = torch.empty(NPARAM, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.wmx_mlp_train_step(_ptr(self.params), _ptr(X), _ptr(Y), X.shape[0], offset,
                                               batch, _ptr(part), _ptr(lpart), _ptr(g), None, None, 0,
                                               0.0, 0, _lib.stream_ptr(stream)))
        return g, lpart.double().sum() / (batch * NOUT)

    def adam_step(self, X, Y, offset, batch, m, v, step, lr, stream=None):
        ng = (batch + 63) // 64
        part = torch.empty(ng * NPARAM, dtype=torch.float32, device=self.device)
        lpart = torch.empty(ng, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.wmx_mlp_train_step(_ptr(self.params), _ptr(X), _ptr(Y), X.shape[0], offset,
                                               batch, _ptr(part), _ptr(lpart), None, _ptr(m), _ptr(v),
                                               step, lr, 1, _lib.stream_ptr(stream)))
        return lpart

    def train(self, gas=None, n_samples=1_000_000, batch=8192, steps=30000, lr=2e-3, lr_min=1e-5,
              warmup=200, seed=0, log_every=0):
        """Train on `n_samples` Newton solves of the method with Adam and a cosine
        schedule.  Returns the history (loss every `log_every` steps, times)."""
        from .gas import Gas
        gas = gas or Gas()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # targets' normalisation from their range over the samples
        X, Y = self.dataset(gas, n_samples, seed)
        lo, hi = torch.aminmax(Y, dim=0)
        lo, hi = lo.double().cpu().numpy(), hi.double().cpu().numpy()
        self.y_center = tuple(float(0.5 * (a + b)) for a, b in zip(lo, hi))
        self.y_scale = tuple(float(max(0.5 * (b - a), 1e-6)) for a, b in zip(lo, hi))
        X, Y = self.dataset(gas, n_samples, seed)
        torch.cuda.synchronize()
        t_data = time.perf_counter() - t0
        if self.params is None:
            self.init_params(seed)
        m = torch.zeros(NPARAM, dtype=torch.float32, device=self.device)
        v = torch.zeros_like(m)
        nb = max(1, n_samples // batch)
        ng = (batch + 63) // 64
        part = torch.empty(ng * NPARAM, dtype=torch.float32, device=self.device)
        lpart = torch.empty(ng, dtype=torch.float32, device=self.device)
        stream = _lib.stream_ptr()
        fn = _lib.lib.wmx_mlp_train_step
        losses = []
        t1 = time.perf_counter()
        for k in range(1, steps + 1):
            if k <= warmup:
                rate = lr * k / warmup
            else:
                c = (k - warmup) / max(1, steps - warmup)
                rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * c))
            off = ((k - 1) % nb) * batch
            _lib.check(fn(_ptr(self.params), _ptr(X), _ptr(Y), n_samples, off, batch, _ptr(part),
                          _ptr(lpart), None, _ptr(m), _ptr(v), k, rate, 1, stream))
            if log_every and k % log_every == 0:
                losses.append((k, float(lpart.double().sum()) / (batch * NOUT)))
        self.pack()
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        _, full = self.grad(X, Y)
        self.trained = True
        self.history = {"losses": losses, "final_mse": float(full), "data_s": t_data,
                        "train_s": t_train, "steps": steps, "batch": batch, "n_samples": n_samples}
        return self.history

    # ------------------------------------------------------------------ inference
    def forward32(self, X, stream=None):
        Y = torch.empty((X.shape[0], 2), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib.wmx_mlp_forward(_ptr(self.params), _ptr(X), X.shape[0], _ptr(Y),
                                            _lib.stream_ptr(stream)))
        return Y

    def eval(self, gas, h, U, T_h, p_h, T_wall=None, hdel=None, mode=0, stream=None):
        """(tau_w, q_w, u+) on point arrays; mode 0 the polished model, mode 1 the raw
        network (diagnostics)."""
        from .backend import asarray
        if self.params is None:
            raise RuntimeError("LearnedWallModel has no parameters: train() or load() first")
        ts = [asarray(x).reshape(-1) for x in (h, U, T_h, p_h)]
        n = max(t.numel() for t in ts)
        ts = [t.expand(n).contiguous() if t.numel() == 1 else t.contiguous() for t in ts]
        hd = None if hdel is None else asarray(hdel).reshape(-1).expand(n).contiguous()
        tau = torch.empty(n, dtype=torch.float64, device=ts[0].device)
        qw, up = torch.empty_like(tau), torch.empty_like(tau)
        nd = _nd(gas, self.base)
        tw = -1.0 if T_wall is None else float(T_wall)
        _lib.check(_lib.lib.wmx_wall_net_eval(C.byref(nd), C.byref(self.desc()), mode, n,
                                              *[_ptr(t) for t in ts[:1]], _ptr(hd),
                                              *[_ptr(t) for t in ts[1:]], tw, _ptr(tau), _ptr(qw),
                                              _ptr(up), _lib.stream_ptr(stream)))
        return tau, qw, up

    def __call__(self, gas, h, U, T_h, p_h, T_wall):
        """(tau_w, q_w): the reference model's callable (wall_model.py:42-61)."""
        tau, qw, _ = self.eval(gas, h, U, T_h, p_h, T_wall)
        shape = tuple(np.broadcast(*[np.empty(np.shape(x)) for x in (h, U, T_h, p_h)]).shape)
        return tau.reshape(shape), qw.reshape(shape)

    def utau(self, h, U, rho_w, mu_w):
        """u_tau from (h, U, rho_w, mu_w) by the Newton model (wall_model.py:19-40)."""
        return self.base.utau(h, U, rho_w, mu_w)

    def host_constants(self, gas):
        return self.base.host_constants(gas)

"""Multiblock explicit finite-volume driver on MI355X (reference `wmles/solver.py`).

`Solver(mesh, gas, scheme, sgs="vreman", cfl=0.8, sgs_kwargs=None, constraints=(),
sources=(), bc_passes=2)` with `init_state`, `apply_bcs`, `rhs`, `max_dt`,
`advance`, `wall_stats` as in the reference.  State = tuple of float64 torch
tensors (5, Ni, Nj, Nk) on the HIP device, ghosts included.  Every call returns new
tensors and leaves its inputs unchanged.  All step arithmetic runs in the HIP
kernels of libwmles_amd.so; `max_dt` returns a 0-d device tensor that `advance`
consumes without a host round trip.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .backend import asarray, device
from .bc import WallModel, is_external
from .mesh import block_desc
from .schemes import KEP, KEEP, MUSCLHLLC, Hybrid
from .sgs import SGSModel, make_sgs


# FP32-bulk plans evaluate the gradient records inside the flux march unless set otherwise
FUSED_GRADIENTS_DEFAULT = True


class ConstantMassFlux:
    """Post-step constraint: shift the volume-averaged rho u_dir of all block
    interiors to `target`, keeping the internal energy (solver.py:16-36).  Inside
    `Solver.advance` it runs as HIP reductions; called directly it applies the same
    kernels through a one-off plan of the solver it is attached to."""

    def __init__(self, direction, target):
        if direction not in (0, 1, 2):
            raise ValueError("direction must be 0, 1 or 2")
        self.dir, self.target = direction, float(target)
        self._solver = None

    def __call__(self, Q, geoms, blocks):
        if self._solver is None:
            raise RuntimeError("ConstantMassFlux is applied by the Solver it was passed to")
        out = tuple(q.clone() for q in self._solver._as_state(Q))
        _lib.check(_lib.lib.wmx_mass_flux(self._solver._plan, _lib.ptr_array(out), _lib.stream_ptr()))
        return out


class _Ctx:
    """Per-solver context handed to the face descriptors (inflow table ids)."""

    def __init__(self):
        self.inflows = []


def wall_model_of(blocks):
    """The one wall-model object the WallModel faces of `blocks` use (None without
    wall-modelled faces).  The wall kernel evaluates a single model per plan."""
    models = []
    for b in blocks:
        for bc in b.faces.values():
            if isinstance(bc, WallModel) and not any(bc.model is m for m in models):
                models.append(bc.model)
    if not models:
        return None
    from .learned_wm import LearnedWallModel
    from .wall_model import EquilibriumWallModel
    for m in models:
        if not isinstance(m, (EquilibriumWallModel, LearnedWallModel)):
            raise NotImplementedError(f"wall model {type(m).__name__} has no HIP implementation")
    first = models[0]
    for m in models[1:]:
        same = (type(m) is type(first) and isinstance(m, EquilibriumWallModel)
                and (m.kappa, m.B, m.n_iter) == (first.kappa, first.B, first.n_iter))
        if not same:
            raise NotImplementedError("all WallModel faces of a solver must share one wall model")
    return first


def numerics_desc(gas, scheme, sgs, cfl, bc_passes, constraints, sources, n_cases=1,
                  precision=_lib.PREC_FP64, wall_model=None):
    nd = _lib.NumericsDesc()
    nd.gamma, nd.R, nd.Pr, nd.Pr_t = gas.gamma, gas.R, gas.Pr, gas.Pr_t
    nd.mu_ref, nd.T_ref, nd.S_suth = gas.mu_ref, gas.T_ref, gas.S_suth
    nd.scheme, nd.central, nd.order, nd.limiter = _lib.SCHEME_KEP, _lib.SCHEME_KEP, 2, 0
    nd.th0, nd.th1 = 0.6, 0.9
    if scheme is not None:
        nd.scheme = scheme.code
        if isinstance(scheme, (KEP, KEEP)):
            nd.order = scheme.order
        elif isinstance(scheme, MUSCLHLLC):
            nd.limiter = 0 if scheme.limiter == "albada" else 1
        elif isinstance(scheme, Hybrid):
            nd.central = scheme.c.code
            nd.order = scheme.c.order
            nd.limiter = 0 if scheme.u.limiter == "albada" else 1
            nd.th0, nd.th1 = float(scheme.th0), float(scheme.th1)
        else:
            raise NotImplementedError(f"scheme {type(scheme).__name__} has no HIP implementation")
    sg = sgs if isinstance(sgs, SGSModel) else make_sgs(None)
    nd.sgs, nd.sgs_c = sg.code, sg.coef
    from .wall_model import EquilibriumWallModel
    wm = wall_model if wall_model is not None else EquilibriumWallModel()
    nd.kappa, nd.B, nd.n_iter = wm.kappa, wm.B, wm.n_iter
    nd.wm_ekb, nd.wm_pr13, nd.wm_pr3, nd.wm_beta = wm.host_constants(gas)
    nd.bc_passes = int(bc_passes)
    nd.cfl = float(cfl)
    nd.precision = precision
    nd.n_sponge = 0
    for src in sources:
        from .smooth_ramp import SmoothRampSponge
        if not isinstance(src, SmoothRampSponge):
            continue                     # a user source: added by the Solver (torch)
        if nd.n_sponge >= 4:
            raise NotImplementedError("at most 4 sponge sources")
        vals = (src.x_start_l, src.x_end_l, src.strength, src.props["U_inf"], src.props["L"])
        for i, v in enumerate(vals):
            nd.sponge[nd.n_sponge][i] = float(v)
        nd.n_sponge += 1
    nd.n_cases = n_cases
    for c in range(16):
        nd.mass_flux_dir[c] = -1
        nd.mass_flux_target[c] = 0.0
    for c, cons in enumerate(constraints):
        nd.mass_flux_dir[c] = cons.dir
        nd.mass_flux_target[c] = cons.target
    return nd


class Solver:
    """solver.py:39-232."""

    def __init__(self, mesh, gas, scheme, sgs="vreman", cfl=0.8, sgs_kwargs=None,
                 constraints=(), sources=(), bc_passes=2, precision="fp64", gauge=None):
        """precision: "fp64" (plain form) or "fp32" (FP32-bulk: FP32 flux, gradient and
        SGS arithmetic; FP64 state, metrics, reductions, stage updates and wall model).
        gauge: (p_ref, T_ref) of the FP32-bulk records; default: the interior means of
        the first state the solver steps."""
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self._gauge = None
        self.mesh, self.blocks, self.gas = mesh, mesh.blocks, gas
        self.geoms = mesh.geoms
        if self.geoms is None:
            raise RuntimeError("call Mesh.finalize() before building the Solver")
        self.scheme, self.cfl = scheme, cfl
        self.sgs = make_sgs(sgs, **(sgs_kwargs or {}))
        self.constraints, self.sources, self.bc_passes = tuple(constraints), tuple(sources), bc_passes
        native = [c for c in self.constraints if isinstance(c, ConstantMassFlux)]
        if len(native) > 1 or (native and self.constraints[0] is not native[0]):
            raise NotImplementedError("one ConstantMassFlux, first in the constraint list, is supported")
        self._native_constraints = native
        self._py_constraints = [c for c in self.constraints if not isinstance(c, ConstantMassFlux)]
        for c in native:
            c._solver = self
        self._dev = device()
        self._ctx = _Ctx()
        self._keep = []
        descs = [block_desc(b, mesh.index, self._ctx, g) for b, g in zip(self.blocks, self.geoms)]
        self._descs = (_lib.BlockDesc * len(descs))(*descs)
        self.wall_model = wall_model_of(self.blocks)
        self._nd = numerics_desc(gas, scheme, self.sgs, cfl, bc_passes, native, self.sources,
                                 precision=_lib.PREC_FP32_BULK if precision == "fp32" else _lib.PREC_FP64,
                                 wall_model=self.wall_model)
        inflows = (_lib.InflowDesc * max(1, len(self._ctx.inflows)))()
        for i, inf in enumerate(self._ctx.inflows):
            inflows[i] = inf.desc(self._dev)
            self._keep.append(inf)
        plan = C.c_void_p()
        _lib.check(_lib.lib.wmx_plan_create(len(descs), self._descs, C.byref(self._nd),
                                             len(self._ctx.inflows), inflows, C.byref(plan)))
        self._plan = plan
        nbytes = C.c_size_t()
        _lib.check(_lib.lib.wmx_plan_workspace_bytes(plan, C.byref(nbytes)))
        self._ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=self._dev)
        _lib.check(_lib.lib.wmx_plan_bind_workspace(plan, C.c_void_p(self._ws.data_ptr()), nbytes.value))
        cnt = C.c_int32()
        _lib.check(_lib.lib.wmx_plan_wall_faces(plan, C.byref(cnt), None))
        rec = (C.c_int32 * (5 * max(1, cnt.value)))()
        _lib.check(_lib.lib.wmx_plan_wall_faces(plan, C.byref(cnt), rec))
        self._wall_faces = [tuple(rec[5 * i:5 * i + 5]) for i in range(cnt.value)]
        self._dt_buf = torch.empty(16, dtype=torch.float64, device=self._dev)
        self._fused = precision == "fp32" and FUSED_GRADIENTS_DEFAULT
        if precision == "fp32":
            self.set_fused_gradients(self._fused)
        # user BCs (no HIP descriptor) and user sources run on the torch side: the sweep
        # is then driven phase by phase from here, and the RK3 stages combine in torch
        from .smooth_ramp import SmoothRampSponge
        self._py_sources = [src for src in self.sources if not isinstance(src, SmoothRampSponge)]
        self._external = any(is_external(bc) for b in self.blocks for bc in b.faces.values())
        self._fallback = self._external or bool(self._py_sources)
        self._phases = []
        if self._external:
            npd = C.c_int32()
            _lib.check(_lib.lib.wmx_plan_sweep_phases(plan, C.byref(npd), None))
            rec = (C.c_int32 * (4 * max(1, npd.value)))()
            _lib.check(_lib.lib.wmx_plan_sweep_phases(plan, C.byref(npd), rec))
            self._phases = [tuple(rec[4 * i:4 * i + 4]) for i in range(npd.value)]
        if gauge is not None:
            self.set_gauge(*gauge)
        from .learned_wm import LearnedWallModel
        if isinstance(self.wall_model, LearnedWallModel):
            self.set_wall_net(self.wall_model)

    def set_wall_net(self, model):
        """Evaluate the wall-modelled faces with a trained LearnedWallModel (None: the
        Newton model)."""
        if model is None:
            _lib.check(_lib.lib.wmx_plan_set_wall_net(self._plan, None))
            self._wall_net = None
            return
        if model.params is None:
            raise RuntimeError("LearnedWallModel has no parameters: train() or load() it before "
                               "building the Solver")
        self._wall_net = model
        _lib.check(_lib.lib.wmx_plan_set_wall_net(self._plan, C.byref(model.desc())))

    def set_gauge(self, p_ref, T_ref):
        """Reference pressure / temperature subtracted in the FP32-bulk records."""
        self._gauge = (float(p_ref), float(T_ref))
        _lib.check(_lib.lib.wmx_plan_set_gauge(self._plan, self._gauge[0], self._gauge[1]))

    def _ensure_gauge(self, Q):
        if self.precision != "fp32" or self._gauge is not None:
            return
        blk, q = self.blocks[0], Q[0]
        w, T = self.gas.cons_to_prim(blk.interior(q))
        self.set_gauge(float(w[4].mean()), float(T.mean()))

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan:
            try:
                torch.cuda.synchronize()
                _lib.lib.wmx_plan_destroy(plan)
            except Exception:
                pass

    # ------------------------------------------------------------------ helpers
    def _as_state(self, Q):
        out = []
        for q, b in zip(Q, self.blocks):
            t = asarray(q)
            if tuple(t.shape) != (5,) + tuple(b.N):
                raise ValueError(f"block {b.name}: state must be (5,)+{b.N}, got {tuple(t.shape)}")
            out.append(t)
        if len(out) != len(self.blocks):
            raise ValueError("one state array per block is required")
        return tuple(out)

    KERNEL_CLASSES = ("flux", "grad", "prim", "fill", "wall", "dt", "massflux", "shell", "inflow", "sponge")

    def profile(self, classes=("flux",)):
        """Start event timing of the named kernel classes (empty: stop)."""
        mask = 0
        for c in classes:
            mask |= 1 << self.KERNEL_CLASSES.index(c)
        _lib.check(_lib.lib.wmx_plan_profile(self._plan, mask))

    def profile_read(self):
        """{class: (total_ms, launches)} of the launches timed since profile()."""
        ms = (C.c_double * len(self.KERNEL_CLASSES))()
        cnt = (C.c_int32 * len(self.KERNEL_CLASSES))()
        _lib.check(_lib.lib.wmx_plan_profile_read(self._plan, ms, cnt))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNEL_CLASSES) if cnt[i]}

    def fill_phases(self):
        """(kernel phases, sweep passes) of the ghost sweep after plan analysis."""
        n, k = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib.wmx_plan_fill_phases(self._plan, C.byref(n)))
        _lib.check(_lib.lib.wmx_plan_fill_passes(self._plan, C.byref(k)))
        return n.value, k.value

    # ------------------------------------------------------------------ init
    def init_state(self, prim_fn):
        """prim_fn(xc) -> [rho, u, v, w, p] at every cell centre, ghosts included
        (solver.py:54-56).  xc is passed as a host NumPy array."""
        out = []
        for g in self.geoms:
            w = prim_fn(g.xc.cpu().numpy())
            w = asarray(np.asarray(w) if not isinstance(w, torch.Tensor) else w)
            out.append(self.gas.prim_to_cons(w).contiguous())
        return tuple(out)

    # ------------------------------------------------------------------ step API
    def apply_bcs(self, Q, geoms=None, time=0.0):
        """Sequential ghost sweep (solver.py:59-73); returns new tensors."""
        out = tuple(q.clone() for q in self._as_state(Q))
        if self._external:
            return self._sweep(list(out), float(time))
        _lib.check(_lib.lib.wmx_apply_bcs(self._plan, _lib.ptr_array(out), float(time), _lib.stream_ptr()))
        return out

    def _sweep(self, Q, time):
        """The sweep with user BCs: runs of HIP phases between the user fills, each user
        fill at its place in the reference's order (solver.py:59-73)."""
        first = 0
        for i, (ext, b, a, side) in enumerate(self._phases + [(1, -1, -1, -1)]):
            if not ext:
                continue
            if i > first:
                _lib.check(_lib.lib.wmx_apply_bcs_phases(self._plan, _lib.ptr_array(Q), first, i, time,
                                                         _lib.stream_ptr()))
            first = i + 1
            if b < 0:
                break
            blk = self.blocks[b]
            q = blk.faces[(a, side)].fill_at_time(Q[b], blk, self.geoms[b], self.gas, a, side, time)
            q = asarray(q)
            if tuple(q.shape) != tuple(Q[b].shape):
                raise ValueError(f"{type(blk.faces[(a, side)]).__name__}.fill returned shape {tuple(q.shape)}")
            Q[b] = q if q.data_ptr() != Q[b].data_ptr() else q.clone()
        return tuple(Q)

    def rhs(self, Q, time=0.0):
        """RHS with zero ghosts (solver.py:97-106,159-160)."""
        if self._fallback:
            return self._rhs_torch(self._as_state(Q), float(time))
        work = tuple(q.clone() for q in self._as_state(Q))
        self._ensure_gauge(work)
        R = tuple(torch.empty_like(q) for q in work)
        _lib.check(_lib.lib.wmx_rhs(self._plan, _lib.ptr_array(work), _lib.ptr_array(R), float(time),
                                    _lib.stream_ptr()))
        return R

    def _rhs_torch(self, Q, time):
        """RHS for plans with user BCs or sources: the sweep (HIP phases and user fills),
        the HIP RHS of the filled state, plus each user source on the interior
        (solver.py:97-106)."""
        Qf = self.apply_bcs(Q, None, time)
        self._ensure_gauge(Qf)
        R = tuple(torch.empty_like(q) for q in Qf)
        _lib.check(_lib.lib.wmx_rhs_filled(self._plan, _lib.ptr_array(Qf), _lib.ptr_array(R), _lib.stream_ptr()))
        for blk, q, r, g in zip(self.blocks, Qf, R, self.geoms):
            for src in self._py_sources:
                add = asarray(src(blk, q, g, self.gas))
                interior = blk.interior(r)
                if tuple(add.shape) != tuple(interior.shape):
                    raise ValueError(f"source {type(src).__name__} returned {tuple(add.shape)}, "
                                     f"expected {tuple(interior.shape)}")
                interior += add
        return R

    def _advance_torch(self, Q, dt, time):
        """SSP-RK3 of solver.py:109-133 with the stage combinations in torch (plans with
        user BCs or sources), then the constraints."""
        if not isinstance(dt, torch.Tensor):
            dt = torch.tensor(float(dt), dtype=torch.float64, device=self._dev)
        dt = dt.reshape(-1)[0] if dt.numel() > 1 else dt.reshape(())
        tdt = float(dt)
        R0 = self._rhs_torch(Q, time)
        Q1 = tuple(q + dt * r for q, r in zip(Q, R0))
        R1 = self._rhs_torch(Q1, time + tdt)
        Q2 = tuple(0.75 * q0 + 0.25 * (q1 + dt * r1) for q0, q1, r1 in zip(Q, Q1, R1))
        R2 = self._rhs_torch(Q2, time + 0.5 * tdt)
        out = tuple((1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * r2) for q0, q2, r2 in zip(Q, Q2, R2))
        if self._native_constraints:
            out = tuple(q.contiguous() for q in out)
            _lib.check(_lib.lib.wmx_mass_flux(self._plan, _lib.ptr_array(out), _lib.stream_ptr()))
        for c in self._py_constraints:
            out = tuple(self._as_state(c(out, self.geoms, self.blocks)))
        return out

    def max_dt(self, Q, time=0.0):
        """CFL time step (solver.py:135-154) as a 0-d device tensor."""
        Q = self._as_state(Q)
        out = torch.empty(16, dtype=torch.float64, device=self._dev)
        _lib.check(_lib.lib.wmx_max_dt(self._plan, _lib.ptr_array(Q), C.c_void_p(out.data_ptr()),
                                       _lib.stream_ptr()))
        return out[0]

    def advance(self, Q, dt, time=0.0):
        """One SSP-RK3 step + constraints (solver.py:123-133,156-157)."""
        Q = self._as_state(Q)
        self._ensure_gauge(Q)
        if self._fallback:
            return self._advance_torch(Q, dt, float(time))
        out = tuple(torch.empty_like(q) for q in Q)
        if isinstance(dt, torch.Tensor):
            d = dt.reshape(-1).to(device=self._dev, dtype=torch.float64)
            if not d.is_contiguous():
                d = d.contiguous()
            rc = _lib.lib.wmx_advance(self._plan, _lib.ptr_array(Q), _lib.ptr_array(out), 0.0,
                                      C.c_void_p(d.data_ptr()), float(time), _lib.stream_ptr())
        else:
            rc = _lib.lib.wmx_advance(self._plan, _lib.ptr_array(Q), _lib.ptr_array(out), float(dt),
                                      None, float(time), _lib.stream_ptr())
        _lib.check(rc)
        for c in self._py_constraints:
            out = tuple(self._as_state(c(out, self.geoms, self.blocks)))
        return out

    def step(self, Q, time=0.0, graph=False):
        """The reference driver's step, max_dt then advance (examples/channel_wmles.py:
        124-125), with dt kept on the device (the CFL min is taken in the first stage's
        primitive pass; same value as max_dt).  Launches go to the caller's stream;
        graph=True instead replays the whole launch sequence as one captured hipGraph
        when the state pointers repeat (the usual `Q, dt = solver.step(Q)` loop alternates
        two buffers).  Returns (Q_new, dt)."""
        Q = self._as_state(Q)
        self._ensure_gauge(Q)
        if self._fallback:            # user BCs / sources: max_dt + the torch-driven stages
            dt = self.max_dt(Q, time)
            return self._advance_torch(Q, dt, float(time)), dt
        if not graph:                 # eager: the caller's stream, no side-stream handshake
            out = tuple(torch.empty_like(q) for q in Q)
            dt = torch.empty(16, dtype=torch.float64, device=self._dev)
            _lib.check(_lib.lib.wmx_step(self._plan, _lib.ptr_array(Q), _lib.ptr_array(out),
                                         C.c_void_p(dt.data_ptr()), float(time), 0, _lib.stream_ptr()))
            for c in self._py_constraints:
                out = tuple(self._as_state(c(out, self.geoms, self.blocks)))
            return out, dt[0]
        # capture needs a stream other than the legacy default one: the caller's own when it
        # is one (no handshake), else a side stream; the dt buffer the graph was captured
        # with is kept
        if not hasattr(self, "_dt_dev"):
            self._dt_dev = torch.empty(16, dtype=torch.float64, device=self._dev)
        cur = torch.cuda.current_stream(self._dev)
        if cur.cuda_stream != 0:
            out = tuple(torch.empty_like(q) for q in Q)
            _lib.check(_lib.lib.wmx_step(self._plan, _lib.ptr_array(Q), _lib.ptr_array(out),
                                         C.c_void_p(self._dt_dev.data_ptr()), float(time), 1,
                                         _lib.stream_ptr(cur)))
            dt = self._dt_dev[0].clone()
            for c in self._py_constraints:
                out = tuple(self._as_state(c(out, self.geoms, self.blocks)))
            return out, dt
        if not hasattr(self, "_stream"):
            self._stream = torch.cuda.Stream(device=self._dev)
        self._stream.wait_stream(cur)
        with torch.cuda.stream(self._stream):
            out = tuple(torch.empty_like(q) for q in Q)
            rc = _lib.lib.wmx_step(self._plan, _lib.ptr_array(Q), _lib.ptr_array(out),
                                   C.c_void_p(self._dt_dev.data_ptr()), float(time), int(bool(graph)),
                                   _lib.stream_ptr(self._stream))
            _lib.check(rc)
            dt = self._dt_dev[0].clone()
        cur.wait_stream(self._stream)
        for q in Q:
            q.record_stream(self._stream)
        for c in self._py_constraints:
            out = tuple(self._as_state(c(out, self.geoms, self.blocks)))
        return out, dt

    def set_fused_gradients(self, on=True):
        """FP32-bulk plans: evaluate the gradient records inside the flux march (the
        default) or in the separate gradient kernel."""
        _lib.check(_lib.lib.wmx_plan_set_fused_gradients(self._plan, 1 if on else 0))
        self._fused = bool(on)

    def set_persistent(self, on=True):
        """Run every RK stage of advance / step as one cooperative (persistent) launch --
        the sweep phases, primitives, gradients and flux tiles behind grid-wide barriers --
        instead of one launch per kernel class (FP32-bulk plans without inflow, sponge,
        learned wall model or user BCs; same results bit for bit)."""
        _lib.check(_lib.lib.wmx_plan_set_persistent(self._plan, 1 if on else 0))
        self.persistent = bool(on)

    def set_tiling(self, ci_len=0, order=1):
        """Flux-kernel tiling: i-planes per marching tile (0: automatic), XCD-interleaved
        order (1) or not (0)."""
        _lib.check(_lib.lib.wmx_plan_set_tiling(self._plan, int(ci_len), int(order)))

    def tiling(self):
        """(i-planes per tile, order, number of tiles) in use."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib.wmx_plan_tiling(self._plan, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    # ------------------------------------------------------------------ diagnostics
    def wall_stats(self, Q):
        """{(block, face): (tau_w, q_w)} for all wall-modelled faces (solver.py:163-179)."""
        Q = self._as_state(Q)
        tot = sum(r[3] * r[4] for r in self._wall_faces)
        tau = torch.empty(max(tot, 1), dtype=torch.float64, device=self._dev)
        qw = torch.empty_like(tau)
        _lib.check(_lib.lib.wmx_wall_stats(self._plan, _lib.ptr_array(Q), C.c_void_p(tau.data_ptr()),
                                           C.c_void_p(qw.data_ptr()), _lib.stream_ptr()))
        tau, qw = tau.cpu().numpy(), qw.cpu().numpy()
        out, off = {}, 0
        for b, a, s, n1, n2 in self._wall_faces:
            cnt = n1 * n2
            key = (self.blocks[b].name, "ijk"[a] + ("min", "max")[s])
            out[key] = (tau[off:off + cnt].reshape(n1, n2), qw[off:off + cnt].reshape(n1, n2))
            off += cnt
        return out

    def wall_loads(self, Q, moment_center=(0.0, 0.0, 0.0)):
        """Integrated fluid-on-wall force and moment per WallModel face
        (solver.py:181-232): tau_w from the HIP wall kernel, integration on the device."""
        center = np.asarray(moment_center, dtype=float)
        if center.shape != (3,) or not np.all(np.isfinite(center)):
            raise ValueError("moment_center must contain three finite coordinates")
        Q = self._as_state(Q)
        stats = self.wall_stats(Q)
        out = {}
        cvec = torch.as_tensor(center, dtype=torch.float64, device=self._dev)
        for blk, q, g in zip(self.blocks, Q, self.geoms):
            w, T = self.gas.cons_to_prim(q)
            for (axis, side), bc in blk.faces.items():
                if not isinstance(bc, WallModel):
                    continue
                ng, n = blk.ng, blk.n[axis]
                fi = ng if side == 0 else ng + n
                mi = ng + bc.n_match - 1 if side == 0 else ng + n - bc.n_match
                ci = ng if side == 0 else ng + n - 1
                Sv = blk.take(g.S[axis], axis, fi)
                area = torch.sqrt((Sv[0] * Sv[0] + Sv[1] * Sv[1]) + Sv[2] * Sv[2])
                nrm = Sv / area
                if side == 0:
                    nrm = -nrm
                vel = blk.take(w[1:4], axis, mi)
                un = (vel[0] * nrm[0] + vel[1] * nrm[1]) + vel[2] * nrm[2]
                ut = vel - un[None] * nrm
                speed = torch.sqrt(((ut[0] ** 2 + ut[1] ** 2) + ut[2] ** 2) + 1e-20)
                tang = ut / speed
                face = "ijk"[axis] + ("min", "max")[side]
                tau = torch.as_tensor(stats[(blk.name, face)][0], device=self._dev)
                pw = blk.take(w[4], axis, ci)
                traction = (pw[None] * nrm + tau[None] * tang) * area[None]
                fc = blk.take(g.xf[axis], axis, fi)
                arm = fc - cvec.reshape(3, 1, 1)
                mom = torch.stack([arm[1] * traction[2] - arm[2] * traction[1],
                                   arm[2] * traction[0] - arm[0] * traction[2],
                                   arm[0] * traction[1] - arm[1] * traction[0]])
                out[(blk.name, face)] = (traction.sum(dim=(1, 2)).cpu().numpy(),
                                         mom.sum(dim=(1, 2)).cpu().numpy())
        return out

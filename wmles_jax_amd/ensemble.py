"""Ensembles of independent cases advanced together (batching form).

`EnsembleSolver(solvers)` merges the blocks of several cases (each set up as an
ordinary `Solver`: its own mesh, BCs, mass-flux target) into one HIP plan.  Every
kernel of the step then covers all blocks of all cases in a single launch (one launch
per kernel class and RK stage for the whole ensemble, instead of one per case), with
per-case CFL time steps and per-case mass-flux reductions, and the step can be
replayed as one captured hipGraph.  Numerics (gas, scheme, SGS, CFL, precision) must be
common to the cases; geometry, sizes, block splits, BCs and targets may differ.
"""
import ctypes as C

import torch

from . import _lib
from .solver import numerics_desc, wall_model_of


def _numerics_key(nd):
    return tuple(getattr(nd, f) for f in ("scheme", "central", "order", "limiter", "th0", "th1", "sgs",
                                          "sgs_c", "cfl", "bc_passes", "precision", "kappa", "B",
                                          "n_iter"))


def merge_block_descs(case_descs):
    """Concatenate the block descriptors of several cases: case c's blocks get
    case_id c and their interface neighbour indices move by the case's block offset.
    Returns (descs, per-case block offsets)."""
    nblk = sum(len(d) for d in case_descs)
    descs = (_lib.BlockDesc * nblk)()
    offsets, k = [], 0
    for c, cd in enumerate(case_descs):
        offsets.append(k)
        for d in cd:
            nd = _lib.BlockDesc()
            C.memmove(C.byref(nd), C.byref(d), C.sizeof(_lib.BlockDesc))
            nd.case_id = c
            for f in range(6):
                if nd.faces[f].kind == _lib.FACE_INTERFACE:
                    nd.faces[f].nb += offsets[c]
                if nd.faces[f].kind == _lib.BC_INFLOW:
                    raise ValueError("time-dependent inflow cases are not batched")
            descs[k] = nd
            k += 1
    return descs, offsets


class EnsembleSolver:
    def __init__(self, solvers):
        solvers = list(solvers)
        if not 1 <= len(solvers) <= 16:
            raise ValueError("an ensemble holds 1..16 cases")
        s0 = solvers[0]
        if any(s._external for s in solvers):
            raise ValueError("ensemble cases must not have user (Python) boundary conditions")
        for s in solvers[1:]:
            same = (s.gas == s0.gas and _numerics_key(s._nd) == _numerics_key(s0._nd)
                    and s.precision == s0.precision and not s.sources and not s0.sources)
            if not same:
                raise ValueError("ensemble cases must share gas, scheme, SGS, CFL, precision and have no sources")
        nblk = sum(len(s.blocks) for s in solvers)
        if nblk > _lib.WMX_MAX_BLOCKS:
            raise ValueError(f"an ensemble plan holds at most {_lib.WMX_MAX_BLOCKS} blocks")
        self.solvers = solvers
        self.precision = s0.precision
        self._dev = s0._dev
        descs, self._offsets = merge_block_descs([s._descs for s in solvers])
        self._descs = descs
        cons = []
        for s in solvers:
            if len(s._native_constraints) != 1 or s._py_constraints:
                raise ValueError("each ensemble case needs exactly one ConstantMassFlux constraint")
            cons.append(s._native_constraints[0])
        self.wall_model = wall_model_of([b for s in solvers for b in s.blocks])
        self._nd = numerics_desc(s0.gas, s0.scheme, s0.sgs, s0.cfl, s0.bc_passes, cons, (),
                                 n_cases=len(solvers),
                                 precision=_lib.PREC_FP32_BULK if self.precision == "fp32" else _lib.PREC_FP64,
                                 wall_model=self.wall_model)
        plan = C.c_void_p()
        _lib.check(_lib.lib.wmx_plan_create(nblk, descs, C.byref(self._nd), 0, None, C.byref(plan)))
        self._plan = plan
        nbytes = C.c_size_t()
        _lib.check(_lib.lib.wmx_plan_workspace_bytes(plan, C.byref(nbytes)))
        self._ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=self._dev)
        _lib.check(_lib.lib.wmx_plan_bind_workspace(plan, C.c_void_p(self._ws.data_ptr()), nbytes.value))
        self._stream = torch.cuda.Stream(device=self._dev)
        self._dt_dev = torch.empty(16, dtype=torch.float64, device=self._dev)
        self._gauge = None            # per-case (p_ref, T_ref) once set
        from .learned_wm import LearnedWallModel
        if isinstance(self.wall_model, LearnedWallModel):
            _lib.check(_lib.lib.wmx_plan_set_wall_net(self._plan, C.byref(self.wall_model.desc())))

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan:
            try:
                torch.cuda.synchronize()
                _lib.lib.wmx_plan_destroy(plan)
            except Exception:
                pass

    @property
    def n_cases(self):
        return len(self.solvers)

    def cells(self):
        return sum(int(b.n[0] * b.n[1] * b.n[2]) for s in self.solvers for b in s.blocks)

    def _flat(self, Qs):
        if len(Qs) != self.n_cases:
            raise ValueError("one state tuple per case")
        flat = []
        for s, Q in zip(self.solvers, Qs):
            flat.extend(s._as_state(Q))
        return flat

    def _split(self, flat):
        out, k = [], 0
        for s in self.solvers:
            out.append(tuple(flat[k:k + len(s.blocks)]))
            k += len(s.blocks)
        return out

    def _ensure_gauge(self, Qs):
        """FP32-bulk records of case c hold p - p_ref(c), T - T_ref(c): each case's gauge
        is its own solver's (set_gauge, or the interior means of its first state)."""
        if self.precision != "fp32" or self._gauge is not None:
            return
        gauges = []
        for c, (s, Q) in enumerate(zip(self.solvers, Qs)):
            s._ensure_gauge(s._as_state(Q))
            gauges.append(s._gauge)
            _lib.check(_lib.lib.wmx_plan_set_case_gauge(self._plan, c, *s._gauge))
        self._gauge = gauges

    def init_state(self, prim_fns):
        return [s.init_state(f) for s, f in zip(self.solvers, prim_fns)]

    def max_dt(self, Qs):
        """Per-case CFL time steps, device tensor (n_cases,)."""
        flat = self._flat(Qs)
        out = torch.empty(16, dtype=torch.float64, device=self._dev)
        _lib.check(_lib.lib.wmx_max_dt(self._plan, _lib.ptr_array(flat), C.c_void_p(out.data_ptr()),
                                       _lib.stream_ptr()))
        return out[:self.n_cases]

    def advance(self, Qs, dts, time=0.0):
        """All cases one SSP-RK3 step with their own dt (device tensor (n_cases,))."""
        flat = self._flat(Qs)
        self._ensure_gauge(Qs)
        d = torch.as_tensor(dts, dtype=torch.float64, device=self._dev).reshape(-1).contiguous()
        if d.numel() != self.n_cases:
            raise ValueError("one dt per case")
        out = [torch.empty_like(q) for q in flat]
        _lib.check(_lib.lib.wmx_advance(self._plan, _lib.ptr_array(flat), _lib.ptr_array(out), 0.0,
                                        C.c_void_p(d.data_ptr()), float(time), _lib.stream_ptr()))
        return self._split(out)

    def step(self, Qs, graph=False):
        """max_dt + advance for every case (caller's stream); graph=True: captured-graph
        replay when the state pointers repeat."""
        flat = self._flat(Qs)
        self._ensure_gauge(Qs)
        if not graph:                 # eager: the caller's stream, no side-stream handshake
            out = [torch.empty_like(q) for q in flat]
            dt = torch.empty(max(16, self.n_cases), dtype=torch.float64, device=self._dev)
            _lib.check(_lib.lib.wmx_step(self._plan, _lib.ptr_array(flat), _lib.ptr_array(out),
                                         C.c_void_p(dt.data_ptr()), 0.0, 0, _lib.stream_ptr()))
            return self._split(out), dt[:self.n_cases]
        cur = torch.cuda.current_stream(self._dev)
        if cur.cuda_stream != 0:      # the caller's capturable stream: no handshake
            out = [torch.empty_like(q) for q in flat]
            _lib.check(_lib.lib.wmx_step(self._plan, _lib.ptr_array(flat), _lib.ptr_array(out),
                                         C.c_void_p(self._dt_dev.data_ptr()), 0.0, 1, _lib.stream_ptr(cur)))
            return self._split(out), self._dt_dev[:self.n_cases].clone()
        self._stream.wait_stream(cur)
        with torch.cuda.stream(self._stream):
            out = [torch.empty_like(q) for q in flat]
            _lib.check(_lib.lib.wmx_step(self._plan, _lib.ptr_array(flat), _lib.ptr_array(out),
                                         C.c_void_p(self._dt_dev.data_ptr()), 0.0, int(bool(graph)),
                                         _lib.stream_ptr(self._stream)))
            dts = self._dt_dev[:self.n_cases].clone()
        cur.wait_stream(self._stream)
        for q in flat:
            q.record_stream(self._stream)
        return self._split(out), dts

    KERNEL_CLASSES = ("flux", "grad", "prim", "fill", "wall", "dt", "massflux", "shell", "inflow", "sponge")

    def profile(self, classes=("flux",)):
        """Start HIP-event timing of the named kernel classes (empty: stop).  Launches
        replayed from a captured graph are not timed."""
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

    def set_fused_gradients(self, on=True):
        """FP32-bulk plans: evaluate the gradient records inside the flux march (the
        default) or in the separate gradient kernel."""
        _lib.check(_lib.lib.wmx_plan_set_fused_gradients(self._plan, 1 if on else 0))

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

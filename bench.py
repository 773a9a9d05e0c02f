"""Throughput of the explicit SSP-RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form both|fp64|fp32]
                    [--workload channel256|ramp512|ensemble16|wm512] [--no-cpu-baseline]
                    [--graph] [--breakdown]

Default workload (channel256): compressible wall-modelled channel, Ma_b 0.3, Re_b 2e4,
one block of 256 x 128 x 128 cells (x, y, z), tanh-clustered y, Hybrid(KEP(4),
MUSCL-HLLC) with the Ducros switch, Vreman SGS, isothermal equilibrium wall model on
both walls, periodic x / z, constant mass flux, CFL 0.8, seeded synthetic initial state
(bulk profile + 5% Gaussian velocity noise).  A step is the reference's driver step:
max_dt + advance (3 RHS + 3 stage updates + constraint), dt kept on the device, run as
Solver.step (the CFL min is taken in the first stage's primitive pass, the same value
as max_dt; eager launches, --graph replays captured hipGraphs instead).  The
timed region holds exactly K steps bracketed by device synchronisations.  Both
arithmetic forms run; the FP32-bulk form is the headline value, the FP64 form and the
FP32-bulk errors (u_tau, mean u(y), q_w against FP64 after the same steps) are reported
beside it.

Other workloads (one JSON line each):
  ramp512     smooth-ramp curvilinear block 512 x 128 x 64, inflow + sponge, FP32-bulk
  ensemble16  16 uneven multiblock channel cases in one plan (one launch per kernel
              class and stage), autotuned tiling, against the
              same cases stepped one after another
  wm512       channel 512 x 256 x 256 at Re_b 1e5 with the learned wall model (trained
              here on the GPU) against the Newton model: wall-kernel time, u_tau error

Printed (one JSON line): Mcell-updates/s, ms/step, the roofline of the dominant
kernel (fused flux / divergence / stage update) from live HIP events (against the
datasheet HBM peak, and beside it the device's measured copy bandwidth), and the CPU
baseline (the FP64 oracle timed on the host on a bounded sample of the same
workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0                      # MI355X HBM3E, datasheet
# Algorithmic bytes per interior cell per RK3 step, as the flux kernels read them.
# State (FP64, both forms): stage 1 reads q and writes q1 (10 words), stages 2 and 3
# read q_k and q0 and write the result (15 words each): 40 words = 320 B.
# Geometry per stage: FP64 plain form 3 owned face vectors (9 words) + xc (3) + vol (1)
# = 104 B; FP32-bulk form float face vectors (36 B) + FP64 xc (24 B) + FP64 vol (8 B)
# = 68 B.  The fused flux kernel moves one third per launch.
B_ALG_STEP = {"fp64": 320.0 + 3 * 104.0, "fp32": 320.0 + 3 * 68.0}   # 632 / 524 B


def make_case(nx, ny, nz, ns):
    import cases
    spec, arrays = cases.channel_spec(1, nx, ny, nz, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", T_wall=300.0, name=f"channel_{nx}x{ny}x{nz}")
    mesh, solver = cases.build(ns, spec, arrays)
    return spec, mesh, solver


def host_threads():
    """Host cores this process may run on, OMP_NUM_THREADS and the BLAS pools NumPy uses."""
    info = {"affinity_cores": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        from threadpoolctl import threadpool_info
        info["blas_threads"] = {p.get("internal_api", "?"): p.get("num_threads") for p in threadpool_info()}
    except Exception as e:                # threadpoolctl absent or failing: say so
        info["blas_threads"] = f"unknown ({type(e).__name__})"
    return info


def cpu_baseline(sample=(32, 128, 128), steps=4):
    """Oracle (NumPy FP64) on a 32x128x128 block of the same numerics.  The oracle's
    array arithmetic is NumPy element-wise code, which runs on one thread whatever the
    BLAS pool size: `cores` is 1, and the host's core count and thread settings are
    reported beside it."""
    import cases
    import oracle
    spec, mesh, s = make_case(*sample, oracle)
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = float(s.max_dt(Q))
        Q = s.advance(Q, dt)
    el = time.perf_counter() - t0
    cells = sample[0] * sample[1] * sample[2]
    return {"value": round(cells * steps / el / 1e6, 5), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port", "host": host_threads(),
            "sample": f"{steps} steps (max_dt + advance) of the same numerics on a "
                      f"{sample[0]}x{sample[1]}x{sample[2]} channel block, NumPy FP64 oracle "
                      f"(element-wise, one thread), {el:.1f} s on the host"}


def profile_record(workload, fused=True):
    """The committed rocprofv3 figures of the flux kernel for `workload` (HBM bytes per
    launch from FETCH_SIZE / WRITE_SIZE, average launch time from the kernel trace), or
    {} when none matches."""
    path = os.path.join(ROOT, "profiles", "flux_traffic.json")
    if not os.path.exists(path):
        return {}
    try:
        recs = json.load(open(path))
    except ValueError:
        return {}
    for rec in recs if isinstance(recs, list) else [recs]:
        if rec.get("workload") == workload and rec.get("fused_gradients", True) == fused:
            return rec
    return {}


def _own_stream():
    """A new stream that first waits for the current one (the work it starts from)."""
    import torch
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    return st


def run_form(wmx, form, size, steps, warmup, breakdown, graph=False, fused=None):
    """Time `steps` steps of one arithmetic form; returns timings, final state, wall stats.
    fused: FP32-bulk kernel form (True: gradient records in the flux march; None: the
    plan's default)."""
    import torch
    import cases
    nx, ny, nz = size
    spec, arrays = cases.channel_spec(1, nx, ny, nz, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", T_wall=300.0, name=f"channel_{nx}x{ny}x{nz}")
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": form})
    if fused is not None and form == "fp32":
        s.set_fused_gradients(fused)
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    # captured-graph steps replay on a stream of the caller's own (the legacy default stream
    # cannot be captured), so the replay needs no side-stream handshake
    with torch.cuda.stream(_own_stream() if graph else torch.cuda.current_stream()):
        for _ in range(warmup):
            Q = s.step(Q, graph=graph)[0]
        torch.cuda.synchronize()
        s.profile(s.KERNEL_CLASSES if breakdown else ("flux",))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            Q, dt = s.step(Q, graph=graph)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(())
    if graph:
        # replayed launches are not event-timed: the kernel times come from eager steps
        s.profile(s.KERNEL_CLASSES if breakdown else ("flux",))
        Qp = Q
        for _ in range(5):
            Qp = s.step(Qp, graph=False)[0]
        torch.cuda.synchronize()
        prof = s.profile_read()
        s.profile(())
    return {"elapsed": el, "prof": prof, "Q": Q, "mesh": mesh, "solver": s, "props": pr}


def form_errors(r64, r32):
    """FP32-bulk against FP64 after the same steps from the same state: u_tau (wall
    mean), plane-mean u(y), q_w (wall mean), all relative."""
    import numpy as np
    blk = r64["mesh"].blocks[0]
    q64, q32 = r64["Q"][0].cpu().numpy(), r32["Q"][0].cpu().numpy()
    u64 = blk.interior(q64[1] / q64[0]).mean(axis=(0, 2))
    u32 = blk.interior(q32[1] / q32[0]).mean(axis=(0, 2))
    st64, st32 = r64["solver"].wall_stats(r64["Q"]), r32["solver"].wall_stats(r32["Q"])
    rho_w = r64["props"]["p0"] / (287.05 * 300.0)
    ut = lambda st: np.sqrt(np.mean([v[0].mean() for v in st.values()]) / rho_w)
    qw = lambda st: np.mean([v[1].mean() for v in st.values()])
    return {"u_tau": float(abs(ut(st32) - ut(st64)) / ut(st64)),
            "mean_u_profile": float(np.abs(u32 - u64).max() / np.abs(u64).max()),
            "q_w": float(abs(qw(st32) - qw(st64)) / abs(qw(st64)))}


FLUX_KERNEL = {True: "k_flux3 (gradient records in the march + face fluxes + divergence + RK stage)",
               False: "k_flux2 (face fluxes + divergence + RK stage; gradient records from k_grad32)",
               None: "k_flux1 (FP64: face fluxes + divergence + RK stage)"}


_COPY_PEAK = []


def copy_peak_gbs():
    """Measured device copy bandwidth (the STREAM copy of this device): 2 GiB of doubles
    copied with torch's vectorised element-wise copy, HIP events around each of 20 copies
    after 5 untimed ones, the best copy; GB/s counts the bytes read plus written."""
    import torch
    if not _COPY_PEAK:
        n = 256 * 1024 * 1024
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        best = float("inf")
        for _ in range(20):
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            y.copy_(x)
            t1.record()
            t1.synchronize()
            best = min(best, t0.elapsed_time(t1) * 1e-3)
        del x, y
        torch.cuda.empty_cache()
        _COPY_PEAK.append(2 * n * 8 / best / 1e9)
    return _COPY_PEAK[0]


def _roofline(avg_launch_s, b_launch, workload, b_cell_step, fused=True):
    """achieved = algorithmic bytes per flux launch / the launch time measured live with
    HIP events on the launch stream.  The committed rocprofv3 trace's average launch
    time for the same workload, and the fraction it gives, are reported beside it."""
    achieved = b_launch / avg_launch_s / 1e9
    prof = profile_record(workload, fused)
    rec = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": prof.get("hbm_bytes_per_launch"),
           "kernel": FLUX_KERNEL[fused],
           "alg_bytes_per_cell_step": b_cell_step,
           "alg_bytes_per_launch": round(b_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4)}
    us = prof.get("rocprof_avg_launch_us")
    if us:
        rec["rocprof_avg_launch_ms"] = round(us / 1e3, 4)
        rec["rocprof_frac"] = round(b_launch / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    if rec["traffic"]:
        rec["traffic_over_alg"] = round(rec["traffic"] / b_launch, 3)
    for k in ("valu_issue_busy", "valu_lane_utilisation"):
        if k in prof:
            rec["rocprof_" + k] = prof[k]
    # the device's measured copy bandwidth beside the datasheet peak
    cp = copy_peak_gbs()
    rec["measured_copy_peak"] = round(cp, 1)
    rec["frac_of_measured_copy_peak"] = round(achieved / cp, 4)
    return rec


def _record(args, value, ms_step, dtype, workload, cfg, roofline):
    return {
        "metric": "Mcell-updates/s per RK3 step, FP32-bulk vs FP64 at stated error; % HBM peak",
        "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": dict({"workload": workload}, **cfg), "roofline": roofline}


def bench_channel(args, wmx):
    nx, ny, nz = args.size
    cells = nx * ny * nz
    forms = ["fp64", "fp32"] if args.form == "both" else [args.form]
    fused = {"auto": None, "on": True, "off": False}[args.fused]
    res = {f: run_form(wmx, f, args.size, args.steps, args.warmup, args.breakdown, args.graph, fused)
           for f in forms}
    head = "fp32" if "fp32" in res else "fp64"
    r = res[head]
    el = r["elapsed"]
    value = cells * args.steps / el / 1e6
    flux_ms, flux_n = r["prof"]["flux"]
    b_launch = B_ALG_STEP[head] / 3.0 * cells
    workload = f"channel {nx}x{ny}x{nz} Hybrid(KEP4+MUSCL-HLLC) Vreman WM {head}"
    rec = _record(args, value, el * 1e3 / args.steps, "f32-bulk+f64" if head == "fp32" else "f64",
                  workload, {"cells": cells, "blocks": 1, "ng": 3,
                             "step": "Solver.step = max_dt + advance (SSP-RK3, 3 RHS, mass-flux "
                                     "constraint; the CFL min taken in the first stage's primitive pass)",
                             "graph": bool(args.graph),
                             "form": head,
                             "fused_gradients": r["solver"]._fused if head == "fp32" else None},
                  _roofline(flux_ms / flux_n / 1e3, b_launch, workload, B_ALG_STEP[head],
                            r["solver"]._fused if head == "fp32" else None))
    rec["step_hbm_frac"] = round(B_ALG_STEP[head] * cells / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4)
    ci, order, ntiles = r["solver"].tiling()
    rec["config"]["tiling"] = {"ci_len": ci, "xcd_order": order, "tiles": ntiles}
    if "fp64" in res and "fp32" in res:
        v64 = cells * args.steps / res["fp64"]["elapsed"] / 1e6
        rec["fp64_value"] = round(v64, 2)
        rec["fp64_ms_per_step"] = round(res["fp64"]["elapsed"] * 1e3 / args.steps, 4)
        rec["fp32_bulk_speedup_vs_fp64"] = round(value / v64, 3)
        rec["fp32_bulk_error_vs_fp64"] = form_errors(res["fp64"], res["fp32"])
    if args.breakdown:
        rec["kernel_ms_per_step"] = {f: {k: round(v[0] / args.steps, 4) for k, v in res[f]["prof"].items()}
                                     for f in res}
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline()
    return rec


def _timed_steps(step, Q, steps, warmup):
    import torch
    for _ in range(warmup):
        Q = step(Q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Q = step(Q)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, Q


def bench_ramp(args, wmx):
    """Smooth-ramp curvilinear block 512 x 128 x 64, FP32-bulk."""
    import torch
    import cases
    nx, ny, nz = 512, 128, 64
    spec, arrays, prim = cases.ramp_spec(wmx, nx, ny, nz)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    Q = s.init_state(prim)
    t = [0.0]

    def step(Q):
        return s.step(Q, time=t[0], graph=False)[0]

    s.profile(("flux",))
    el, Q = _timed_steps(step, Q, args.steps, args.warmup)
    prof = s.profile_read()
    s.profile(())
    cells = nx * ny * nz
    finite = all(bool(torch.isfinite(q).all()) for q in Q)
    workload = f"smooth ramp {nx}x{ny}x{nz} Hybrid(KEP4+MUSCL-HLLC) Vreman WM inflow+sponge fp32"
    flux_ms, flux_n = prof["flux"]
    rec = _record(args, cells * args.steps / el / 1e6, el * 1e3 / args.steps, "f32-bulk+f64", workload,
                  {"cells": cells, "blocks": 1, "ng": 3, "form": "fp32"},
                  _roofline(flux_ms / flux_n / 1e3, B_ALG_STEP["fp32"] / 3.0 * cells, workload,
                            B_ALG_STEP["fp32"]))
    rec["state_finite"] = finite
    return rec


def bench_ensemble(args, wmx):
    """16 uneven multiblock channel cases: one merged plan against sequential cases."""
    import torch
    import cases
    from wmles_jax_amd import autotune
    specs = cases.ensemble_specs(16)
    solvers, fns = [], []
    for spec, arrays, pr, seed in specs:
        _, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
        solvers.append(s)
        fns.append(lambda xc, pr=pr, seed=seed: cases.channel_prim(xc, pr, seed))
    cells = sum(int(b.n[0] * b.n[1] * b.n[2]) for s in solvers for b in s.blocks)
    nblk = sum(len(s.blocks) for s in solvers)
    ens = wmx.EnsembleSolver(solvers)
    Qs0 = ens.init_state(fns)
    if args.ens_forms == "eager":                # (no autotuner launches: the plan's own tiling)
        best, timings = ens.tiling()[:2], {}
    else:
        best, timings = autotune.tune(ens, Qs0, use_cache=False)

    def ens_step(Qs):
        return ens.step(Qs, graph=args.graph)[0]

    with torch.cuda.stream(_own_stream() if args.graph else torch.cuda.current_stream()):
        el_e, Qe = _timed_steps(ens_step, Qs0, args.steps, args.warmup)
    # flux launch time from events on eager (non-replayed) steps of the same plan
    ens.profile(("flux",))
    Qp = Qs0
    for _ in range(5):
        Qp = ens.step(Qp, graph=False)[0]
    prof = ens.profile_read()
    ens.profile(())
    # the same cases stepped one after another (the same step form per case), with the
    # ensemble's tiling (each case keeps its own FP32-bulk gauge, as in the ensemble) so
    # the results can be compared bitwise
    for s in solvers:
        s.set_tiling(*best)

    def seq_step(Qs):
        return [s.step(Q, graph=args.graph)[0] for s, Q in zip(solvers, Qs)]

    workload = f"ensemble of 16 uneven channel cases ({nblk} blocks) KEP2 WALE WM fp32"
    flux_ms, flux_n = prof["flux"]
    rec = _record(args, cells * args.steps / el_e / 1e6, el_e * 1e3 / args.steps, "f32-bulk+f64", workload,
                  {"cells": cells, "cases": 16, "blocks": nblk, "form": "fp32",
                   "tiling": {"ci_len": best[0], "xcd_order": best[1]}},
                  _roofline(flux_ms / flux_n / 1e3, B_ALG_STEP["fp32"] / 3.0 * cells, workload,
                            B_ALG_STEP["fp32"]))
    rec["state_finite"] = all(bool(torch.isfinite(q).all()) for Q in Qe for q in Q)
    rec["autotune_flux_ms"] = {k: round(v, 4) for k, v in timings.items()}
    if args.ens_forms == "eager":
        # the merged plan's eager steps only (a kernel trace of the ensemble launches alone)
        return rec
    el_s, Qseq = _timed_steps(seq_step, Qs0, args.steps, args.warmup)
    same = all(torch.equal(a, b) for Qa, Qb in zip(Qe, Qseq) for a, b in zip(Qa, Qb))
    # the other launch structures of the same plan: captured-graph replay, and one
    # persistent (cooperative) launch per RK stage
    # (replayed on a stream of the caller's own, so no side-stream handshake: the caller's
    # stream must not be the legacy default stream to be captured)
    with torch.cuda.stream(_own_stream()):
        el_g, Qg = _timed_steps(lambda Qs: ens.step(Qs, graph=True)[0], Qs0, args.steps, args.warmup)
    pers = wmx.EnsembleSolver(solvers)
    pers.set_tiling(*best)
    pers.set_persistent(True)
    el_p, Qpe = _timed_steps(lambda Qs: pers.step(Qs)[0], Qs0, args.steps, args.warmup)
    same_g = all(torch.equal(a, b) for Qa, Qb in zip(Qe, Qg) for a, b in zip(Qa, Qb))
    same_p = all(torch.equal(a, b) for Qa, Qb in zip(Qe, Qpe) for a, b in zip(Qa, Qb))
    rec["sequential_value"] = round(cells * args.steps / el_s / 1e6, 2)
    rec["ensemble_speedup_vs_sequential"] = round(el_s / el_e, 3)
    rec["ensemble_equals_sequential"] = same
    rec["launch_forms_ms_per_step"] = {"eager": round(el_e * 1e3 / args.steps, 4),
                                       "graph_replay": round(el_g * 1e3 / args.steps, 4),
                                       "persistent_stage": round(el_p * 1e3 / args.steps, 4)}
    rec["graph_equals_eager"] = same_g
    rec["persistent_equals_eager"] = same_p
    return rec


def bench_wm(args, wmx):
    """Learned wall model in the 512 x 256 x 256 channel at Re_b 1e5."""
    import numpy as np
    import torch
    import cases
    nx, ny, nz = 512, 256, 256
    model = wmx.LearnedWallModel()
    hist = model.train()
    # the 256-cell tanh grid resolves the wall (Re_h of the first cell is about 12, below the
    # trained range); the model samples the 10th cell off the wall (Re_h about 400)
    n_match = 10
    spec, arrays = cases.channel_spec(1, nx, ny, nz, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", T_wall=300.0, Re_b=1e5, name=f"channel_{nx}x{ny}x{nz}",
                                      n_match=n_match)
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": "fp32"})
    pr = spec["props"]
    Q0 = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    rho_w = pr["p0"] / (287.05 * 300.0)

    def run(net):
        s.set_wall_net(net)
        st0 = s.wall_stats(Q0)
        s.profile(("wall", "flux"))
        el, Q = _timed_steps(lambda Q: s.step(Q, graph=args.graph)[0], Q0, args.steps, args.warmup)
        prof = s.profile_read()
        s.profile(())
        if args.graph:                     # replayed launches are not event-timed
            s.profile(("wall", "flux"))
            Qp = Q
            for _ in range(5):
                Qp = s.step(Qp, graph=False)[0]
            torch.cuda.synchronize()
            prof = s.profile_read()
            s.profile(())
        return el, Q, prof, st0

    el_n, Qn, prof_n, st_n = run(None)
    el_l, Ql, prof_l, st_l = run(model)
    # The network's own accuracy (raw outputs, mode 1) and the shipped model's (network
    # start + Newton polish, mode 0) against the Newton model, on 1e5 points spread over
    # the trained range (log-uniform Re_h in [1e2, 1e6], T_h/T_w in [0.8, 1.25]).
    gas = wmx.Gas()
    rng = np.random.default_rng(7)
    npts, T_w, p_w, U = 100_000, 300.0, 1e5, 10.0
    Reh = np.exp(rng.uniform(np.log(1e2), np.log(1e6), npts))
    rho_w, mu_w = p_w / (gas.R * T_w), float(gas.mu(T_w))
    h_t = torch.tensor(Reh * mu_w / (U * rho_w), device="cuda")
    Th_t = torch.tensor(T_w * rng.uniform(0.8, 1.25, npts), device="cuda")
    tau_n, qw_n = wmx.EquilibriumWallModel()(gas, h_t, U, Th_t, p_w, T_w)
    net_err = {}
    for name, mode in (("raw_network", 1), ("polished", 0)):
        tau_m, qw_m, _ = model.eval(gas, h_t, U, Th_t, p_w, T_wall=T_w, mode=mode)
        ut_err = torch.abs(torch.sqrt(tau_m / tau_n) - 1.0)
        net_err[name] = {"u_tau_rel_err_max": float(ut_err.max()), "u_tau_rel_err_mean": float(ut_err.mean()),
                         "q_w_err_max_over_max_q_w": float(torch.abs(qw_m - qw_n).max() / torch.abs(qw_n).max())}
    n_faces = sum(int(v[0].size) for v in st_l.values())
    model.fallbacks.zero_()
    s.wall_stats(Q0)                      # one evaluation of every wall face
    fb = int(model.fallbacks.item())
    # same state: per-face tau_w of the two models; after the steps: wall-mean u_tau
    d_tau = max(float(np.max(np.abs(st_l[k][0] - st_n[k][0]) / np.abs(st_n[k][0]))) for k in st_n)
    st_ln = s.wall_stats(Ql)
    s.set_wall_net(None)
    st_nn = s.wall_stats(Qn)
    ut = lambda st: float(np.sqrt(np.mean([float(v[0].mean()) for v in st.values()]) / rho_w))
    cells = nx * ny * nz
    workload = f"channel {nx}x{ny}x{nz} Re_b 1e5 Hybrid(KEP4+MUSCL-HLLC) Vreman learned WM fp32"
    flux_ms, flux_n = prof_l["flux"]
    rec = _record(args, cells * args.steps / el_l / 1e6, el_l * 1e3 / args.steps, "f32-bulk+f64", workload,
                  {"cells": cells, "blocks": 1, "form": "fp32", "wall_faces": n_faces,
                   "network": "3-64-64-2 ISRU, BF16 MFMA inference, 3 Newton polish steps",
                   "n_match": n_match},
                  _roofline(flux_ms / flux_n / 1e3, B_ALG_STEP["fp32"] / 3.0 * cells, workload,
                            B_ALG_STEP["fp32"]))
    # FP32-bulk plans evaluate the Newton faces in another launch's head (no launch of
    # their own); the learned model keeps its wall launch.  Both step times are reported.
    per = lambda pr: round(pr["wall"][0] / pr["wall"][1], 5) if "wall" in pr else None
    rec["wall_model"] = {
        "newton_ms_per_launch": per(prof_n),
        # the head of the primitive launch (fused form) or of the gradient launch (unfused)
        "newton_in_shared_launch": "wall" not in prof_n,
        "learned_ms_per_launch": per(prof_l),
        "step_speedup_learned_vs_newton": round(el_n / el_l, 3),
        "fallback_fraction": round(fb / max(1, n_faces), 6), "n_match": n_match,
        "accuracy_vs_newton_1e5_points": net_err,
        "tau_w_max_rel_err_same_state": d_tau,
        "u_tau_rel_err_after_steps": abs(ut(st_ln) - ut(st_nn)) / ut(st_nn),
        "newton_value": round(cells * args.steps / el_n / 1e6, 2),
        "train_s": round(hist["train_s"], 3), "data_s": round(hist["data_s"], 3),
        "train_steps": hist["steps"], "train_samples": hist["n_samples"], "final_mse": hist["final_mse"]}
    return rec


def bench_copy(args, wmx):
    """Plain device copy of 512 MiB (64 Mi doubles): the known-size kernel the HBM
    traffic counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) are calibrated against."""
    import torch
    n = 64 * 1024 * 1024
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    el, _ = _timed_steps(lambda _: y.copy_(x), None, args.steps, args.warmup)
    gbs = 2 * n * 8 * args.steps / el / 1e9
    return {"metric": "device copy GB/s (counter calibration)", "value": round(gbs, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(el * 1e3 / args.steps, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "copy 512 MiB", "bytes_read_per_launch": n * 8,
                       "bytes_written_per_launch": n * 8}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--form", default="both", choices=["both", "fp64", "fp32"])
    ap.add_argument("--workload", default="channel256",
                    choices=["channel256", "ramp512", "ensemble16", "wm512", "copy"])
    ap.add_argument("--size", type=int, nargs=3, default=(256, 128, 128))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="time every kernel class (extra events)")
    ap.add_argument("--fused", default="auto", choices=["auto", "on", "off"],
                    help="FP32-bulk: gradient records in the flux march (on) or a separate kernel (off)")
    ap.add_argument("--ens-forms", default="all", choices=["all", "eager"],
                    help="ensemble16: also time the sequential cases, graph replay and persistent "
                         "stages (all), or the merged plan's eager steps only (eager)")
    ap.add_argument("--graph", action="store_true",
                    help="replay captured hipGraphs for Solver.step (eager launches otherwise)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"channel256": 200, "ramp512": 200, "ensemble16": 100, "wm512": 20,
                      "copy": 20}[args.workload]

    import torch
    import wmles_jax_amd as wmx
    torch.cuda.set_device(0)
    fn = {"channel256": bench_channel, "ramp512": bench_ramp, "ensemble16": bench_ensemble,
          "wm512": bench_wm, "copy": bench_copy}[args.workload]
    rec = fn(args, wmx)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()

"""Throughput of the explicit SSP-RK3 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp64] [--no-cpu-baseline]

Workload: compressible wall-modelled channel, Ma_b 0.3, Re_b 2e4, one block of
256 x 128 x 128 cells (x, y, z), tanh-clustered y, Hybrid(KEP(4), MUSCL-HLLC) with the
Ducros switch, Vreman SGS, isothermal equilibrium wall model on both walls, periodic
x / z, constant mass flux, CFL 0.8, seeded synthetic initial state (bulk profile +
5% Gaussian velocity noise).  A step is the reference's driver step: max_dt + advance
(3 RHS + 3 stage updates + constraint), dt kept on the device.  The timed region
holds exactly K steps bracketed by device synchronisations.

Printed (one JSON line): Mcell-updates/s, ms/step, the roofline of the dominant
kernel (fused flux / divergence / stage update) from live HIP events, and the CPU
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
# algorithmic bytes per interior cell per RK3 step: FP64 state 40 words + FP64
# geometry 13 words x 3 stages; the fused flux kernel moves one third per launch.
B_ALG_STEP = {"fp64": 632.0, "fp32": 632.0}   # S kept in FP64 in both forms


def make_case(nx, ny, nz, ns):
    import cases
    spec, arrays = cases.channel_spec(1, nx, ny, nz, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", T_wall=300.0, name=f"channel_{nx}x{ny}x{nz}")
    mesh, solver = cases.build(ns, spec, arrays)
    return spec, mesh, solver


def cpu_baseline(sample=(32, 128, 128), steps=1):
    """Oracle (NumPy FP64, single-threaded) on a 32x128x128 block of the same numerics."""
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
            "kind": "port",
            "sample": f"{steps} step(s) (max_dt + advance) of the same numerics on a "
                      f"{sample[0]}x{sample[1]}x{sample[2]} channel block, NumPy FP64 oracle, "
                      f"{el:.1f} s on the host"}


def traffic_from_profile(workload):
    path = os.path.join(ROOT, "profiles", "flux_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path))
    except ValueError:
        return None
    if rec.get("workload") != workload:
        return None
    return rec.get("hbm_bytes_per_launch")


def run_form(wmx, form, size, steps, warmup, breakdown):
    """Time `steps` steps of one arithmetic form; returns timings, final state, wall stats."""
    import torch
    import cases
    nx, ny, nz = size
    spec, arrays = cases.channel_spec(1, nx, ny, nz, {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
                                      "vreman", T_wall=300.0, name=f"channel_{nx}x{ny}x{nz}")
    mesh, s = cases.build(wmx, spec, arrays, solver_kwargs={"precision": form})
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    for _ in range(warmup):
        Q = s.advance(Q, s.max_dt(Q))
    torch.cuda.synchronize()
    s.profile(s.KERNEL_CLASSES if breakdown else ("flux",))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = s.max_dt(Q)
        Q = s.advance(Q, dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--form", default="both", choices=["both", "fp64", "fp32"])
    ap.add_argument("--size", type=int, nargs=3, default=(256, 128, 128))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="time every kernel class (extra events)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import wmles_jax_amd as wmx
    torch.cuda.set_device(0)
    nx, ny, nz = args.size
    cells = nx * ny * nz
    forms = ["fp64", "fp32"] if args.form == "both" else [args.form]
    res = {f: run_form(wmx, f, args.size, args.steps, args.warmup, args.breakdown) for f in forms}
    head = "fp32" if "fp32" in res else "fp64"
    r = res[head]
    el = r["elapsed"]
    ms_step = el * 1e3 / args.steps
    value = cells * args.steps / el / 1e6
    flux_ms, flux_n = r["prof"]["flux"]
    avg_flux_s = flux_ms / flux_n / 1e3
    b_launch = B_ALG_STEP[head] / 3.0 * cells
    achieved = b_launch / avg_flux_s / 1e9
    workload = f"channel {nx}x{ny}x{nz} Hybrid(KEP4+MUSCL-HLLC) Vreman WM {head}"
    rec = {
        "metric": "Mcell-updates/s per RK3 step, FP32-bulk vs FP64 at stated error; % HBM peak",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-bulk+f64" if head == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "blocks": 1, "ng": 3,
                   "step": "max_dt + advance (SSP-RK3, 3 RHS, mass-flux constraint)",
                   "form": head},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(workload),
                     "kernel": "k_flux (fused face fluxes + divergence + RK stage)",
                     "alg_bytes_per_launch": b_launch, "avg_launch_ms": round(avg_flux_s * 1e3, 4)},
        "step_hbm_frac": round(B_ALG_STEP[head] * cells / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
    }
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
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()

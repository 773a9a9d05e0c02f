"""Flux-kernel tiling autotuner (project-written).

Times a few steps of a solver (or ensemble) for each candidate tiling -- i-planes per
marching tile (0: the plan's makespan model) x XCD-interleaved tile order -- with HIP
events around the flux kernel,
keeps the fastest, and caches the choice per (device, form, block sizes) in a JSON file
(default ~/.cache/wmles_jax_amd/tiling.json, override with WMX_TUNE_CACHE).
"""
import json
import os

import torch

CANDIDATES = ((0, 1), (0, 0), (8, 1), (12, 1), (16, 1), (24, 1), (32, 1))


def _cache_path():
    return os.environ.get("WMX_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "wmles_jax_amd", "tiling.json"))


def _key(obj):
    solvers = getattr(obj, "solvers", [obj])
    dims = [tuple(b.n) for s in solvers for b in s.blocks]
    scheme = type(solvers[0].scheme).__name__
    return f"{torch.cuda.get_device_name()}|{solvers[0].precision}|{scheme}|{dims}"


def tune(obj, Q, steps=4, candidates=CANDIDATES, use_cache=True, rounds=2, warmup=10):
    """Pick the fastest tiling for `obj` (Solver or EnsembleSolver) stepping state Q: after
    `warmup` untimed steps (clocks and caches settle), every candidate is timed over
    `steps` steps in each of `rounds` rounds (candidate order reversed in every other
    round), and keeps its best round.  Q is not changed.
    Returns ((ci_len, order), {candidate: ms per flux launch})."""
    path = _cache_path()
    key = _key(obj)
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except ValueError:
            cache = {}
    if use_cache and key in cache:
        ci, order = cache[key]["best"]
        obj.set_tiling(ci, order)
        return (ci, order), cache[key]["timings"]
    timings = {}
    obj.set_tiling(*candidates[0])
    Qw = Q
    for _ in range(warmup):
        Qw, _ = obj.step(Qw, graph=False)
    torch.cuda.synchronize()
    for rnd in range(rounds):
        for ci, order in (candidates if rnd % 2 == 0 else candidates[::-1]):
            obj.set_tiling(ci, order)
            Qw = Q
            Qw, _ = obj.step(Qw, graph=False)        # the candidate's first launches
            torch.cuda.synchronize()
            prof = obj if hasattr(obj, "profile") else None
            if prof is not None:
                prof.profile(("flux",))
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(steps):
                Qw, _ = obj.step(Qw, graph=False)
            t1.record()
            torch.cuda.synchronize()
            if prof is not None:
                ms, n = prof.profile_read().get("flux", (0.0, 0))
                prof.profile(())
                t = ms / max(n, 1)
            else:
                t = t0.elapsed_time(t1) / steps
            k = f"{ci},{order}"
            timings[k] = min(timings.get(k, t), t)
    best = min(timings, key=timings.get)
    ci, order = (int(v) for v in best.split(","))
    obj.set_tiling(ci, order)
    if use_cache:
        cache[key] = {"best": [ci, order], "timings": timings}
        os.makedirs(os.path.dirname(path), exist_ok=True)
        json.dump(cache, open(path, "w"), indent=1)
    return (ci, order), timings

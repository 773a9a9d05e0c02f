"""Write the golden vectors under tests/golden/ from the reference solver itself.

Run in the build container only (the reference tree is not present elsewhere):

    python tests/golden/make_golden.py [/path/to/reference]

The reference is imported with its NumPy backend (WMLES_BACKEND=numpy, float64) and
used for exactly the parts that no closed form or invariant pins to 1e-12 or to the
bit: the composed multiblock RHS with mirror BCs, Hybrid, SGS, wall flux and
sources (reference `wmles/solver.py:76-106`); the edge/corner contents of the
sequential ghost sweep (`wmles/solver.py:59-73`); the synthetic inflow draws
(`wmles/smooth_ramp.py:116-197`); and one advance / max_dt / wall_stats on the same
inputs.  The files hold inputs and outputs only (data, no source).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))          # tests/ -> cases.py
import cases  # noqa: E402


def _load_reference(path):
    os.environ["WMLES_BACKEND"] = "numpy"
    sys.path.insert(0, path)
    import wmles                      # noqa: F401
    import wmles.smooth_ramp          # noqa: F401
    import wmles.schemes              # noqa: F401
    if wmles.HAS_JAX:
        raise RuntimeError("expected the NumPy backend")
    return wmles


def _store_case(out, key, ns, spec, arrays, state_fn, advance=True, unique=False, times=(0.0,)):
    mesh, solver = cases.build(ns, spec, arrays)
    out[f"{key}__spec"] = np.array(json.dumps(spec))
    for k, v in arrays.items():
        out[f"{key}__{k}"] = np.asarray(v)
    Q = state_fn(mesh, solver)
    for i, q in enumerate(Q):
        out[f"{key}__Qin_{i}"] = np.asarray(q)
    for it, t in enumerate(times):
        Qbc = solver.apply_bcs(Q, solver.geoms, t)
        for i, q in enumerate(Qbc):
            out[f"{key}__Qbc{it}_{i}"] = np.asarray(q)
    if unique:
        return
    for it, t in enumerate(times):
        R = solver.rhs(Q, t)
        for i, (blk, r) in enumerate(zip(mesh.blocks, R)):
            out[f"{key}__R{it}_{i}"] = np.asarray(blk.interior(np.asarray(r)))
    dt = float(solver.max_dt(Q))
    out[f"{key}__dt"] = np.array(dt)
    if advance:
        Qa = solver.advance(Q, dt, times[-1])
        for i, q in enumerate(Qa):
            out[f"{key}__Qadv_{i}"] = np.asarray(q)
    st = solver.wall_stats(Q)
    for (bname, face), (tau, qw) in st.items():
        out[f"{key}__tau_{bname}_{face}"] = np.asarray(tau)
        out[f"{key}__qw_{bname}_{face}"] = np.asarray(qw)


def _mb_state(seed, noise_amp=0.0):
    def fn(mesh, solver):
        Q = solver.init_state(cases.prim_fn)
        Q = tuple(np.asarray(q) for q in Q)
        if noise_amp:
            Q = cases.interior_noise(mesh.blocks, Q, seed + 100, noise_amp)
        return cases.noisy_ghost_state(mesh.blocks, Q, seed)
    return fn


def _mixed_spec(n=(6, 5, 4)):
    """Single skewed block, z periodic, and a different non-interface BC on each
    of the other four faces."""
    ix, iy, iz = (np.arange(k + 1) for k in n)
    arrays = {"nodes_all": cases.nodes_from_index(ix, iy, iz, n)}
    prim_inf = [1.2, 40.0, -5.0, 3.0, 1.0e5]
    spec = {"name": "mixed", "blocks": ["all"], "ng": 3,
            "connect": [["all", "kmin", "all", "kmax", [0.0, 0.0, -cases.L[2]]]],
            "bcs": [["all", "imin", {"kind": "Extrapolate"}],
                    ["all", "imax", {"kind": "CharacteristicFarfield", "prim": prim_inf}],
                    ["all", "jmin", {"kind": "NoSlipWall", "T_wall": 280.0}],
                    ["all", "jmax", {"kind": "Dirichlet", "prim": [1.1, 30.0, 2.0, -1.0, 0.98e5]}]],
            "scheme": {"kind": "KEP", "order": 2}, "sgs": "wale", "sgs_kwargs": {},
            "cfl": 0.6, "constraints": [], "sources": [], "gas": {}}
    return spec, arrays


def _supersonic_state(seed):
    def fn(mesh, solver):
        def prim(xc):
            base = cases.prim_fn(xc)
            base[1] = base[1] + 480.0
            return base
        Q = tuple(np.asarray(q) for q in solver.init_state(prim))
        Q = cases.interior_noise(mesh.blocks, Q, seed + 100, 0.05)
        return cases.noisy_ghost_state(mesh.blocks, Q, seed)
    return fn


def _ramp_case(ref):
    props = ref.smooth_ramp.flow_properties(0.2, 667000.0)
    Lr = props["L"]
    nx, ny, nz = 8, 6, 4
    x = np.linspace(-0.1, 1.2, nx + 1)
    xi = np.clip(x, 0.0, 1.0)
    wall = 0.22 * (1.0 - (10 * xi ** 3 - 15 * xi ** 4 + 6 * xi ** 5))
    eta = np.linspace(0.0, 1.0, ny + 1)
    st = 1.0 + np.tanh(1.5 * (eta - 1.0)) / np.tanh(1.5)
    y = wall[:, None] + (0.74 - wall[:, None]) * st[None, :]
    z = np.linspace(0.0, 0.288, nz + 1)
    shp = (nx + 1, ny + 1, nz + 1)
    nodes = np.stack([np.broadcast_to(x[:, None, None], shp), np.broadcast_to(y[:, :, None], shp),
                      np.broadcast_to(z[None, None, :], shp)]) * Lr
    yt = np.linspace(0.22, 0.74, 40)
    U = np.minimum(((yt - 0.22) / 0.028) ** (1.0 / 7.0), 1.0)
    prof = np.stack([yt, U, 0.01 * U, 0.08 ** 2 + 0 * yt, 0.06 ** 2 + 0 * yt,
                     0.05 ** 2 + 0 * yt, -0.3 * 0.08 * 0.06 + 0 * yt], axis=1)
    prim_inf = [props["rho_inf"], props["U_inf"], 0.0, 0.0, props["p_inf"]]
    rp = {k: float(v) for k, v in props.items() if k != "gas"}
    spec = {"name": "ramp", "blocks": ["ramp"], "ng": 3,
            "connect": [["ramp", "kmin", "ramp", "kmax", [0.0, 0.0, -0.288 * Lr]]],
            "bcs": [["ramp", "imin", {"kind": "SmoothRampInflow", "length_scales": [0.1, 0.02, 0.03],
                                      "seed": 3, "n_modes": 16, "span_cells": nz}],
                    ["ramp", "imax", {"kind": "CharacteristicFarfield", "prim": prim_inf}],
                    ["ramp", "jmin", {"kind": "WallModel", "T_wall": None, "n_match": 1}],
                    ["ramp", "jmax", {"kind": "SlipWall"}]],
            "scheme": {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
            "sgs": "vreman", "sgs_kwargs": {}, "cfl": 0.5, "constraints": [],
            "sources": [{"kind": "SmoothRampSponge", "x_start_l": 0.5, "x_end_l": 1.2,
                         "strength": 1.0}],
            "gas": {}, "ramp_props": rp}
    arrays = {"nodes_ramp": nodes, "inflow_profile": prof}

    def state(mesh, solver):
        def prim(xc):
            x_l = xc[0] / Lr
            xi = np.clip(x_l, 0.0, 1.0)
            wl = 0.22 * (1.0 - (10 * xi ** 3 - 15 * xi ** 4 + 6 * xi ** 5))
            yd = xc[1] / Lr - wl + 0.22
            u = props["U_inf"] * np.interp(yd, prof[:, 0], prof[:, 1])
            zero = 0.0 * u
            return np.stack([props["rho_inf"] + zero, u, 0.01 * u, zero, props["p_inf"] + zero])
        Q = tuple(np.asarray(q) for q in solver.init_state(prim))
        Q = cases.interior_noise(mesh.blocks, Q, 77, 0.05)
        return cases.noisy_ghost_state(mesh.blocks, Q, 78)
    return spec, arrays, state, (0.0, 0.05 * Lr / props["U_inf"])


def main(ref_path):
    ref = _load_reference(ref_path)
    n = (6, 4, 4)
    kep4 = {"kind": "KEP", "order": 4}
    hyb = {"kind": "Hybrid", "central": kep4, "upwind": {"kind": "MUSCLHLLC"}, "th0": 0.6, "th1": 0.9}
    wm_iso = {"kind": "WallModel", "T_wall": 300.0, "n_match": 1}
    wm_adi = {"kind": "WallModel", "T_wall": None, "n_match": 2}
    mflux = [{"kind": "ConstantMassFlux", "direction": 0, "target": 1.2 * 25.0}]

    g1 = {}
    plan = [
        ("single_kep4_slip", "single", kep4, {"kind": "SlipWall"}, "vreman", 0.0, []),
        ("split_kep4_slip", "split", kep4, {"kind": "SlipWall"}, "vreman", 0.0, []),
        ("flip_kep4_slip", "flip", kep4, {"kind": "SlipWall"}, "vreman", 0.0, mflux),
        ("rotcyc_kep4_slip", "rot_cyc", kep4, {"kind": "SlipWall"}, "vreman", 0.0, []),
        ("split_hybrid_slip", "split", hyb, {"kind": "SlipWall"}, "vreman", 0.05, mflux),
        ("split_kep4_wm", "split", kep4, wm_iso, "vreman", 0.02, mflux),
        ("rotcyc_hybrid_wm", "rot_cyc", hyb, wm_iso, "smagorinsky", 0.05, []),
        ("single_muscl_noslip", "single", {"kind": "MUSCLHLLC", "limiter": "albada"},
         {"kind": "NoSlipWall", "T_wall": None}, "smagorinsky", 0.05, []),
        ("single_keep4_wmadi", "single", {"kind": "KEEP", "order": 4}, wm_adi, None, 0.02, []),
        ("flip_kep2_wale", "flip", {"kind": "KEP", "order": 2}, wm_iso, "wale", 0.03, mflux),
    ]
    for seed, (key, kind, sc, ybc, sgs, amp, cons) in enumerate(plan):
        spec, arrays = cases.multiblock_spec(kind, n, sc, ybc, sgs=sgs, constraints=cons, name=key)
        _store_case(g1, key, ref, spec, arrays, _mb_state(seed, amp))
    spec, arrays = _mixed_spec()
    _store_case(g1, "mixed_kep2_wale", ref, spec, arrays, _mb_state(20, 0.02))
    spec, arrays = cases.multiblock_spec("single", n, {"kind": "MUSCLHLLC", "limiter": "minmod"},
                                         periodic_y=True, sgs="vreman", name="supersonic_minmod")
    _store_case(g1, "supersonic_minmod", ref, spec, arrays, _supersonic_state(21))
    np.savez_compressed(os.path.join(HERE, "g1_multiblock.npz"), **g1)

    g2 = {}

    def ids(mesh, solver):
        out, base = [], 1.0
        for blk in mesh.blocks:
            cnt = 5 * int(np.prod(blk.N))
            q = (base + np.arange(cnt, dtype=float)).reshape((5,) + blk.N)
            q[0] += 1.0e6                      # keep "rho" positive and away from 0
            q[4] += 1.0e9                      # keep "p" positive for the mirror fills
            out.append(q)
            base += cnt
        return tuple(out)
    for kind in ("split", "flip", "rot_cyc"):
        spec, arrays = cases.multiblock_spec(kind, (6, 5, 4), kep4, periodic_y=True, name=f"ids_{kind}")
        _store_case(g2, f"ids_{kind}", ref, spec, arrays, ids, unique=True)
        spec, arrays = cases.multiblock_spec(kind, (6, 5, 4), kep4, {"kind": "SlipWall"},
                                             name=f"ids_{kind}_slip")
        _store_case(g2, f"ids_{kind}_slip", ref, spec, arrays, ids, unique=True)
    np.savez_compressed(os.path.join(HERE, "g2_unique_ids.npz"), **g2)

    g4 = {}
    spec, arrays = cases.channel_spec(2, 12, 6, 4, {"kind": "KEP", "order": 2}, "wale",
                                      name="channel_small")
    mesh, solver = cases.build(ref, spec, arrays)
    props = spec["props"]
    g4["channel__spec"] = np.array(json.dumps(spec))
    for k, v in arrays.items():
        g4[f"channel__{k}"] = v
    Q = tuple(np.asarray(q) for q in solver.init_state(lambda xc: cases.channel_prim(xc, props, 0)))
    for i, q in enumerate(Q):
        g4[f"channel__Qin_{i}"] = q
    t = 0.0
    for step in range(2):
        dt = float(solver.max_dt(Q))
        Q = solver.advance(Q, dt, t)
        t += dt
        g4[f"channel__dt{step}"] = np.array(dt)
        for i, q in enumerate(Q):
            g4[f"channel__Q{step}_{i}"] = np.asarray(q)
    for (bname, face), (tau, qw) in solver.wall_stats(Q).items():
        g4[f"channel__tau_{bname}_{face}"] = np.asarray(tau)
        g4[f"channel__qw_{bname}_{face}"] = np.asarray(qw)
    np.savez_compressed(os.path.join(HERE, "g4_channel.npz"), **g4)

    g5 = {}
    spec, arrays, state, times = _ramp_case(ref)
    _store_case(g5, "ramp", ref, spec, arrays, state, times=times)
    np.savez_compressed(os.path.join(HERE, "g5_ramp.npz"), **g5)
    print("golden vectors written")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

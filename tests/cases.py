"""Case specifications shared by the golden-vector writer, the oracle tests and the
GPU parity tests.

A spec is plain data (JSON-serialisable dict + named float64 arrays) so the same
case can be built through any namespace that exposes the reference API
(`Block`, `Mesh`, BC classes, schemes, `Solver`, `ConstantMassFlux`): the
reference package (golden writer only), `oracle`, or `wmles_jax_amd`.

The multiblock fixtures restate the reference test geometry and state
(`tests/test_multiblock.py:10-74`): a skewed periodic-compatible grid and a smooth
analytic primitive field.
"""
import json

import numpy as np

L = (1.0, 0.6, 0.5)


def nodes_from_index(ix, iy, iz, n, skew=0.04):
    """Skewed curvilinear periodic-compatible nodes (tests/test_multiblock.py:14-20)."""
    X, Y, Z = np.meshgrid(ix / n[0] * L[0], iy / n[1] * L[1], iz / n[2] * L[2], indexing="ij")
    x = X + skew * np.sin(2 * np.pi * Y / L[1]) * np.sin(2 * np.pi * Z / L[2])
    y = Y + skew * 0.5 * np.sin(2 * np.pi * X / L[0])
    z = Z + skew * 0.5 * np.sin(2 * np.pi * X / L[0]) * np.cos(2 * np.pi * Y / L[1])
    return np.stack([x, y, z])


def prim_fn(xc):
    """Smooth analytic primitive state (tests/test_multiblock.py:23-31)."""
    x, y, z = xc
    k = 2 * np.pi
    rho = 1.2 * (1 + 0.05 * np.sin(k * x / L[0]) * np.cos(k * z / L[2]))
    u = 60 * np.sin(k * y / L[1]) + 20 * np.cos(k * z / L[2])
    v = 15 * np.sin(k * x / L[0]) * np.sin(k * z / L[2])
    w = 10 * np.cos(k * x / L[0]) + 5 * np.sin(k * y / L[1])
    p = 1e5 * (1 + 0.03 * np.cos(k * x / L[0]) * np.sin(k * y / L[1]))
    return np.stack([rho, u, v, w, p])


# ----------------------------------------------------------------------------- specs
def multiblock_spec(kind, n=(12, 10, 8), scheme=None, ybc=None, periodic_y=False,
                    sgs="vreman", cfl=0.5, constraints=(), name=None):
    """Single block or 2 blocks split in x with block B re-oriented
    (tests/test_multiblock.py:34-74).  x and z periodic; y faces get `ybc`."""
    scheme = scheme or {"kind": "KEP", "order": 4}
    ybc = ybc or {"kind": "SlipWall"}
    ix, iy, iz = np.arange(n[0] + 1), np.arange(n[1] + 1), np.arange(n[2] + 1)
    arrays, blocks, connect, bcs = {}, [], [], []
    if kind == "single":
        arrays["nodes_all"] = nodes_from_index(ix, iy, iz, n)
        blocks = ["all"]
        connect.append(["all", "imin", "all", "imax", [-L[0], 0.0, 0.0]])
        zpair = {"all": ("kmin", "kmax")}
        yfaces = {"all": ("jmin", "jmax")}
    else:
        h = n[0] // 2
        nA = nodes_from_index(ix[:h + 1], iy, iz, n)
        nB = nodes_from_index(ix[h:], iy, iz, n)
        if kind == "rot_cyc":
            nB = np.ascontiguousarray(np.transpose(nB, (0, 2, 3, 1)))
        elif kind == "flip":
            nB = np.ascontiguousarray(nB[:, ::-1, ::-1, :])
        arrays["nodes_A"], arrays["nodes_B"] = nA, nB
        blocks = ["A", "B"]
        fx = {"split": ("imin", "imax"), "flip": ("imax", "imin"), "rot_cyc": ("kmin", "kmax")}[kind]
        fz = {"split": ("kmin", "kmax"), "flip": ("kmin", "kmax"), "rot_cyc": ("jmin", "jmax")}[kind]
        fy = {"split": ("jmin", "jmax"), "flip": ("jmax", "jmin"), "rot_cyc": ("imin", "imax")}[kind]
        connect.append(["A", "imax", "B", fx[0], [0.0, 0.0, 0.0]])
        connect.append(["B", fx[1], "A", "imin", [L[0], 0.0, 0.0]])
        zpair = {"A": ("kmin", "kmax"), "B": fz}
        yfaces = {"A": ("jmin", "jmax"), "B": fy}
    for b in blocks:
        connect.append([b, zpair[b][0], b, zpair[b][1], [0.0, 0.0, -L[2]]])
        if periodic_y:
            connect.append([b, yfaces[b][0], b, yfaces[b][1], [0.0, -L[1], 0.0]])
        else:
            for f in yfaces[b]:
                bcs.append([b, f, dict(ybc)])
    spec = {"name": name or f"mb_{kind}", "blocks": blocks, "ng": 3, "connect": connect,
            "bcs": bcs, "scheme": scheme, "sgs": sgs, "sgs_kwargs": {}, "cfl": cfl,
            "constraints": list(constraints), "sources": [], "gas": {}}
    return spec, arrays


def channel_nodes(x0, x1, nx, ny, nz, d, Lz, beta=2.0):
    """Channel block: uniform x and z, tanh wall-clustered y in [0, 2d]
    (examples/channel_wmles.py:26-31 pattern)."""
    x = np.linspace(x0, x1, nx + 1)
    eta = np.linspace(0, 1, ny + 1)
    y = d * (1 + np.tanh(beta * (2 * eta - 1)) / np.tanh(beta))
    z = np.linspace(0, Lz, nz + 1)
    return np.stack(np.meshgrid(x, y, z, indexing="ij"))


def channel_props(Re_b=2.0e4, Ma_b=0.3, T0=300.0, rho0=1.2):
    """Bulk-flow scales of the compressible channel with default gas constants."""
    gamma, R = 1.4, 287.05
    mu_ref, T_ref, S = 1.716e-5, 273.15, 110.4
    mu0 = mu_ref * (T0 / T_ref) ** 1.5 * (T_ref + S) / (T0 + S)
    Ub = Ma_b * np.sqrt(gamma * R * T0)
    d = Re_b * (mu0 / rho0) / Ub
    return dict(T0=T0, rho0=rho0, p0=rho0 * R * T0, Ub=Ub, d=d, mu0=mu0, Re_b=Re_b)


def channel_spec(nblocks_x, nx, ny, nz, scheme, sgs, T_wall=300.0, Re_b=2.0e4, beta=2.0,
                 name="channel", mass_flux=True, splits=None, n_match=1, rho0=1.2):
    """Wall-modelled channel 2 pi d x 2 d x pi d, `nblocks_x` blocks in x (periodic
    x/z, WallModel on jmin/jmax sampling the `n_match`-th cell off the wall, constant mass
    flux).  `nx` is the total x count; `rho0` sets the bulk density (p0 = rho0 R T0)."""
    pr = channel_props(Re_b=Re_b, rho0=rho0)
    d = pr["d"]
    Lx, Lz = 2 * np.pi * d, np.pi * d
    if splits is None:
        cuts = [round(i * nx / nblocks_x) for i in range(nblocks_x + 1)]
    else:
        cuts = list(splits)
    xs = np.linspace(0.0, Lx, nx + 1)
    arrays, blocks, connect, bcs = {}, [], [], []
    for b in range(nblocks_x):
        nm = f"B{b}"
        blocks.append(nm)
        arrays[f"nodes_{nm}"] = channel_nodes(xs[cuts[b]], xs[cuts[b + 1]], cuts[b + 1] - cuts[b],
                                              ny, nz, d, Lz, beta)
    for b in range(nblocks_x):
        nxt = (b + 1) % nblocks_x
        sh = [0.0, 0.0, 0.0] if nxt != 0 else [Lx, 0.0, 0.0]
        if nblocks_x == 1:
            connect.append([blocks[b], "imax", blocks[b], "imin", [Lx, 0.0, 0.0]])
        else:
            connect.append([blocks[b], "imax", blocks[nxt], "imin", sh])
    for nm in blocks:
        connect.append([nm, "kmin", nm, "kmax", [0.0, 0.0, -Lz]])
        for f in ("jmin", "jmax"):
            bcs.append([nm, f, {"kind": "WallModel", "T_wall": T_wall, "n_match": n_match}])
    cons = [{"kind": "ConstantMassFlux", "direction": 0, "target": pr["rho0"] * pr["Ub"]}] \
        if mass_flux else []
    spec = {"name": name, "blocks": blocks, "ng": 3, "connect": connect, "bcs": bcs,
            "scheme": scheme, "sgs": sgs, "sgs_kwargs": {}, "cfl": 0.8,
            "constraints": cons, "sources": [], "gas": {}, "props": pr}
    return spec, arrays


def channel_prim(xc, props, seed=0, noise=0.05, interior=None):
    """Uniform rho, T; u = U_b [1/2 (3/2)(1 - eta^2) + 1/2 (8/7) max(1-|eta|, 1e-3)^(1/7)]
    plus seeded Gaussian (u, v, w) noise of `noise` U_b."""
    pr = props
    d, Ub = pr["d"], pr["Ub"]
    eta = xc[1] / d - 1.0
    u = Ub * (0.5 * 1.5 * (1.0 - eta ** 2)
              + 0.5 * (8.0 / 7.0) * np.maximum(1.0 - np.abs(eta), 1e-3) ** (1.0 / 7.0))
    vel = np.stack([u, np.zeros_like(u), np.zeros_like(u)])
    if noise:
        rng = np.random.default_rng(seed)
        vel = vel + noise * Ub * rng.standard_normal(vel.shape)
    rho = pr["rho0"] + 0.0 * u
    p = pr["p0"] + 0.0 * u
    return np.stack([rho, vel[0], vel[1], vel[2], p])


def ramp_spec(ns, nx=512, ny=128, nz=64, first_dy=1e-3, n_modes=64, seed=0, cfl=0.5):
    """Synthetic smooth-ramp block: x/L in [-0.76, 4], wall from y/L = 0.22 to 0 by a
    quintic over x/L in [0, 1], top at y/L = 0.74, one-sided tanh wall-normal stretching
    with first spacing `first_dy` L at the inflow, periodic span 0.288 L; Ma 0.2,
    Re_L 667000; 1/7-power inflow with Fourier-mode fluctuations of 8% U rms
    (uu = vv = ww = 0.08^2, uv = 0);
    wall-modelled (adiabatic) bottom, slip top, characteristic outflow, sponge over the
    last 1 L.  Returns (spec, arrays, prim_fn)."""
    props = _resolve(ns, "flow_properties")(0.2, 667000.0)
    Lr = props["L"]
    x = np.linspace(-0.76, 4.0, nx + 1)
    xi = np.clip(x, 0.0, 1.0)
    wall = 0.22 * (1.0 - (10 * xi ** 3 - 15 * xi ** 4 + 6 * xi ** 5))
    eta = np.linspace(0.0, 1.0, ny + 1)

    def st(beta):
        return 1.0 + np.tanh(beta * (eta - 1.0)) / np.tanh(beta)

    lo, hi = 1e-3, 20.0                         # first spacing decreases with beta
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if st(mid)[1] * 0.52 > first_dy:
            lo = mid
        else:
            hi = mid
    s_eta = st(0.5 * (lo + hi))
    y = wall[:, None] + (0.74 - wall[:, None]) * s_eta[None, :]
    z = np.linspace(0.0, 0.288, nz + 1)
    shp = (nx + 1, ny + 1, nz + 1)
    nodes = np.stack([np.broadcast_to(x[:, None, None], shp), np.broadcast_to(y[:, :, None], shp),
                      np.broadcast_to(z[None, None, :], shp)]) * Lr
    yt = 0.22 + 0.52 * np.linspace(0.0, 1.0, 64) ** 2
    U = np.minimum(((yt - 0.22) / 0.028) ** (1.0 / 7.0), 1.0)
    a = 0.08 ** 2
    prof = np.stack([yt, U, 0.0 * yt, a + 0 * yt, a + 0 * yt, a + 0 * yt, 0.0 * yt], axis=1)
    prim_inf = [props["rho_inf"], props["U_inf"], 0.0, 0.0, props["p_inf"]]
    rp = {k: float(v) for k, v in props.items() if k != "gas"}
    spec = {"name": f"ramp_{nx}x{ny}x{nz}", "blocks": ["ramp"], "ng": 3,
            "connect": [["ramp", "kmin", "ramp", "kmax", [0.0, 0.0, -0.288 * Lr]]],
            "bcs": [["ramp", "imin", {"kind": "SmoothRampInflow", "length_scales": [0.1, 0.02, 0.03],
                                      "seed": seed, "n_modes": n_modes, "span_cells": nz}],
                    ["ramp", "imax", {"kind": "CharacteristicFarfield", "prim": prim_inf}],
                    ["ramp", "jmin", {"kind": "WallModel", "T_wall": None, "n_match": 1}],
                    ["ramp", "jmax", {"kind": "SlipWall"}]],
            "scheme": {"kind": "Hybrid", "central": {"kind": "KEP", "order": 4}},
            "sgs": "vreman", "sgs_kwargs": {}, "cfl": cfl, "constraints": [],
            "sources": [{"kind": "SmoothRampSponge", "x_start_l": 3.0, "x_end_l": 4.0,
                         "strength": 1.0}],
            "gas": {}, "ramp_props": rp}
    arrays = {"nodes_ramp": nodes, "inflow_profile": prof}

    def prim(xc):
        x_l = xc[0] / Lr
        xx = np.clip(x_l, 0.0, 1.0)
        wl = 0.22 * (1.0 - (10 * xx ** 3 - 15 * xx ** 4 + 6 * xx ** 5))
        yd = xc[1] / Lr - wl + 0.22
        u = props["U_inf"] * np.interp(yd, prof[:, 0], prof[:, 1])
        zero = 0.0 * u
        return np.stack([props["rho_inf"] + zero, u, zero, zero, props["p_inf"] + zero])

    return spec, arrays, prim


def ensemble_specs(n_cases=16, scheme=None, sgs="wale"):
    """`n_cases` channel cases, case s drawn from `default_rng(s)`: Re_b log-uniform in
    [5e3, 4e4], (nx, ny, nz) multiples of 16 in [64, 192] x [48, 96] x [32, 64], each
    split in x into 2-4 blocks at random unequal cuts (every block at least 8 cells).
    Numerics default to the two-block channel's: KEP(2), WALE.
    Returns [(spec, arrays, props, seed)]."""
    scheme = scheme or {"kind": "KEP", "order": 2}
    out = []
    for c in range(n_cases):
        rng = np.random.default_rng(c)
        Re_b = float(np.exp(rng.uniform(np.log(5e3), np.log(4e4))))
        nx = int(rng.integers(4, 13)) * 16
        ny = int(rng.integers(3, 7)) * 16
        nz = int(rng.integers(2, 5)) * 16
        nb = int(rng.integers(2, 5))
        w = rng.uniform(0.5, 1.5, nb)
        cuts = np.concatenate([[0.0], np.cumsum(w) / w.sum()]) * nx
        cuts = [0] + [int(round(v)) for v in cuts[1:-1]] + [nx]
        for i in range(1, nb):
            cuts[i] = min(max(cuts[i], cuts[i - 1] + 8), nx - 8 * (nb - i))
        spec, arrays = channel_spec(nb, nx, ny, nz, scheme, sgs, Re_b=Re_b, splits=cuts,
                                    name=f"case{c}_{nx}x{ny}x{nz}_{nb}blk")
        out.append((spec, arrays, spec["props"], c))
    return out


def dumps(spec):
    return json.dumps(spec)


def loads(text):
    return json.loads(str(text))


# ----------------------------------------------------------------------------- build
def _resolve(ns, name):
    if hasattr(ns, name):
        return getattr(ns, name)
    for sub in ("schemes", "bc", "solver", "wall_model", "smooth_ramp", "ramp"):
        mod = getattr(ns, sub, None)
        if mod is not None and hasattr(mod, name):
            return getattr(mod, name)
    raise AttributeError(f"{ns.__name__} has no {name}")


def make_scheme(ns, sc):
    kind = sc["kind"]
    if kind in ("KEP", "KEEP"):
        return _resolve(ns, kind)(sc.get("order", 4))
    if kind == "MUSCLHLLC":
        return _resolve(ns, "MUSCLHLLC")(sc.get("limiter", "albada"))
    if kind == "Hybrid":
        c = make_scheme(ns, sc["central"]) if sc.get("central") else None
        u = make_scheme(ns, sc["upwind"]) if sc.get("upwind") else None
        return _resolve(ns, "Hybrid")(c, u, sc.get("th0", 0.6), sc.get("th1", 0.9))
    raise ValueError(kind)


def make_bc(ns, bc, props=None, arrays=None):
    kind = bc["kind"]
    if kind == "Object":                      # a BC object built by the test itself
        return bc["obj"]
    if kind == "WallModel":
        wm = bc["model"] if "model" in bc else _resolve(ns, "EquilibriumWallModel")(**bc.get("wm", {}))
        return _resolve(ns, "WallModel")(wm, T_wall=bc.get("T_wall"), n_match=bc.get("n_match", 1))
    if kind == "NoSlipWall":
        return _resolve(ns, "NoSlipWall")(bc.get("T_wall"))
    if kind in ("Dirichlet", "CharacteristicFarfield"):
        return _resolve(ns, kind)(np.asarray(bc["prim"], float))
    if kind in ("SlipWall", "Extrapolate"):
        return _resolve(ns, kind)()
    if kind == "SmoothRampInflow":
        return _resolve(ns, "SmoothRampInflow")(
            props, arrays["inflow_profile"], tuple(bc["length_scales"]), bc["seed"],
            bc["n_modes"], bc["span_cells"])
    raise ValueError(kind)


def build(ns, spec, arrays, solver_kwargs=None):
    """Build (mesh, solver) for `spec` through namespace `ns`."""
    props = spec.get("ramp_props")
    if props is not None:
        props = dict(props)
        props["gas"] = _resolve(ns, "Gas")()
    blocks = {nm: _resolve(ns, "Block")(nm, arrays[f"nodes_{nm}"], spec.get("ng", 3))
              for nm in spec["blocks"]}
    mesh = _resolve(ns, "Mesh")([blocks[nm] for nm in spec["blocks"]])
    for a, fa, b, fb, sh in spec["connect"]:
        mesh.connect(blocks[a], fa, blocks[b], fb, shift=tuple(sh))
    for nm, face, bc in spec["bcs"]:
        blocks[nm].set_bc(face, make_bc(ns, bc, props, arrays))
    mesh.finalize()
    gas = _resolve(ns, "Gas")(**spec.get("gas", {}))
    cons = []
    for c in spec.get("constraints", []):
        cons.append(_resolve(ns, "ConstantMassFlux")(c["direction"], c["target"]))
    srcs = []
    for s in spec.get("sources", []):
        srcs.append(_resolve(ns, "SmoothRampSponge")(props, s["x_start_l"], s["x_end_l"],
                                                     s.get("strength", 1.0)))
    kw = dict(sgs=spec.get("sgs", "vreman"), cfl=spec.get("cfl", 0.8),
              sgs_kwargs=spec.get("sgs_kwargs") or None, constraints=cons, sources=srcs)
    kw.update(solver_kwargs or {})
    solver = _resolve(ns, "Solver")(mesh, gas, make_scheme(ns, spec["scheme"]), **kw)
    return mesh, solver


def noisy_ghost_state(blocks, Q, seed):
    """Perturb every ghost cell by a seeded relative factor so that a fill that
    failed to overwrite (or wrongly read) a ghost value shows up."""
    rng = np.random.default_rng(seed)
    out = []
    for blk, q in zip(blocks, Q):
        q = np.array(q, dtype=float, copy=True)
        mask = np.ones(q.shape[1:], bool)
        ng = blk.ng
        mask[ng:ng + blk.n[0], ng:ng + blk.n[1], ng:ng + blk.n[2]] = False
        fac = 1.0 + 0.01 * rng.standard_normal(q.shape)
        q = np.where(mask[None], q * fac, q)
        out.append(q)
    return tuple(out)


def interior_noise(blocks, Q, seed, amp=0.02):
    """Seeded velocity noise on interior cells (keeps rho and p positive) so that the
    Ducros sensor and the Hybrid switch take values across [0, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for blk, q in zip(blocks, Q):
        q = np.array(q, dtype=float, copy=True)
        ng = blk.ng
        it = (slice(ng, ng + blk.n[0]), slice(ng, ng + blk.n[1]), slice(ng, ng + blk.n[2]))
        rho = q[0][it]
        for c in (1, 2, 3):
            dm = amp * 80.0 * rho * rng.standard_normal(rho.shape)
            E = q[4][it]
            q[4][it] = E + (2 * q[c][it] * dm + dm * dm) / (2 * rho)
            q[c][it] = q[c][it] + dm
        out.append(q)
    return tuple(out)

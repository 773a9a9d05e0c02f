"""Plot3D files and NMF mesh assembly: formats and topology on the CPU; multiblock RHS and
free-stream preservation of assembled meshes on the GPU (the reference's
tests/test_plot3d_nmf.py:50-275; the L5 airfoil grid and its NMF are data files of the
reference's tests, kept here as fixtures)."""
import os
import tempfile
import textwrap

import numpy as np
import pytest

import cases
from wmles_jax_amd import nmf as NM
from wmles_jax_amd import plot3d as P3
from wmles_jax_amd.bc import Dirichlet, NoSlipWall, SlipWall

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _nmf(text):
    with tempfile.NamedTemporaryFile("w", suffix=".nmf", delete=False) as f:
        f.write(textwrap.dedent(text))
    try:
        return NM.read_nmf(f.name)
    finally:
        os.unlink(f.name)


def test_plot3d_round_trip():
    rng = np.random.default_rng(0)
    nodes = [rng.normal(size=(3, 4, 5, 6)), rng.normal(size=(3, 3, 7, 4))]
    ibl = [rng.integers(0, 2, size=X.shape[1:]) for X in nodes]
    with tempfile.TemporaryDirectory() as d:
        for formatted, dt, en, use_ibl in ((True, "f8", "<", False), (True, "f8", "<", True),
                                           (False, "f8", "<", False), (False, "f8", ">", True),
                                           (False, "f4", "<", False)):
            path = os.path.join(d, "g.p3d")
            P3.write_plot3d(path, nodes, formatted=formatted, iblank=ibl if use_ibl else None, dtype=dt, endian=en)
            g = P3.read_plot3d(path)
            assert g.formatted == formatted and g.ndim == 3
            tol = dict(rtol=1e-6, atol=1e-6) if dt == "f4" else dict(rtol=0, atol=1e-13)
            for a, b in zip(g.nodes, nodes):
                assert a.shape == b.shape and np.allclose(a, b, **tol)
            if use_ibl:
                assert all(np.array_equal(a, b) for a, b in zip(g.iblank, ibl))
            else:
                assert g.iblank is None


def test_plot3d_fortran_layout():
    """An unformatted file built by hand: big-endian record markers, i fastest."""
    import struct
    X = np.arange(3 * 2 * 3 * 1, dtype=np.float64).reshape(3, 2, 3, 1)
    rec = lambda b: struct.pack(">i", len(b)) + b + struct.pack(">i", len(b))
    body = b"".join(X[c].ravel(order="F").astype(">f8").tobytes() for c in range(3))
    raw = rec(struct.pack(">i", 1)) + rec(struct.pack(">3i", 2, 3, 1)) + rec(body)
    with tempfile.NamedTemporaryFile(suffix=".x", delete=False) as f:
        f.write(raw)
    try:
        g = P3.read_plot3d(f.name)
    finally:
        os.unlink(f.name)
    assert not g.formatted and np.array_equal(g.nodes[0], X)


def test_plot3d_2d_and_extrude():
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "g2.p2dfmt")
        ni, nj = 4, 3
        x, y = np.meshgrid(np.linspace(0, 1, ni), np.linspace(0, 2, nj), indexing="ij")
        with open(path, "w") as f:
            f.write(f"1\n{ni} {nj}\n")
            # Fortran D exponents
            f.write(" ".join(f"{v:.15E}".replace("E", "D")
                             for v in np.concatenate([x.ravel(order="F"), y.ravel(order="F")])) + "\n")
        g = P3.read_plot3d(path)
        assert g.ndim == 2 and g.nodes[0].shape == (3, ni, nj, 1)
        assert np.allclose(g.nodes[0][0, :, :, 0], x) and np.allclose(g.nodes[0][1, :, :, 0], y)
        ext = NM.extrude_2d(g.nodes, 5, 0.3)
        assert ext[0].shape == (3, ni, nj, 6) and np.isclose(ext[0][2].max(), 0.3)


def test_nmf_parse():
    nmf = _nmf("""\
        # header \\
        # Block#   IDIM   JDIM   KDIM \\
        2 \\
        1 5 6 7 \\
        2 7 5 6 \\
        # Type B1 F1 S1 E1 S2 E2 B2 F2 S1 E1 S2 E2 Swap
        WALL      1 1  1 5  1 6
        Symmetry-Y 1 2  1 5  1 6
        ONE_TO_ONE 1 4  1 6  1 7   2 5  6 1  1 7  TRUE
        """)
    assert nmf.dims == [(5, 6, 7), (7, 5, 6)]
    assert [p.type for p in nmf.patches] == ["wall", "symmetry_y", "one_to_one"]
    o = nmf.patches[2]
    assert o.block == 0 and o.face == 4 and o.partner == dict(block=1, face=5, ranges=((6, 1), (1, 7)), swap=True)
    with pytest.raises(NotImplementedError):
        _nmf("1\n1 4 4 4\ncollapsed 1 1 1 4 1 4\n")


def test_l5_airfoil_nmf_topology():
    """The L5 NMF: 2 blocks; after respan the wake block is cut at the trailing-edge line
    (its jmax face meets the airfoil block's jmax and jmin faces), giving b1_000, b1_001, b2."""
    nmf = NM.read_nmf(os.path.join(GOLDEN, "nak_a_ff_500c_2zn_L5.nmf"))
    assert nmf.dims == [(2, 43, 57), (2, 65, 29)]
    kinds = sorted({p.type for p in nmf.patches})
    assert kinds == ["farfield_riem", "one_to_one", "symmetry_y", "viscous_solid"]
    fake = [np.zeros((3,) + d) for d in nmf.dims]
    for X in fake:
        X[0, 1] = 1.0                                   # a pure translation along axis 0
    _, nmf4 = NM.respan(fake, nmf, ncells=4, length=0.1)
    assert nmf4.dims == [(5, 43, 57), (5, 65, 29)]
    cuts = NM._cut_planes(nmf4)
    assert cuts[0] == [set(), set(), {28}] and cuts[1] == [set(), set(), set()]


def _box(x, y, z, nx, ny, nz):
    X = np.stack(np.meshgrid(np.linspace(*x, nx + 1), np.linspace(*y, ny + 1), np.linspace(*z, nz + 1),
                             indexing="ij"))
    P = X.copy()
    P[0] += 0.04 * np.sin(3 * X[1]) * np.sin(2 * X[2])
    P[1] += 0.03 * np.sin(2 * X[0]) * np.cos(X[2])
    P[2] += 0.03 * np.sin(2 * X[0]) * np.sin(3 * X[1])
    return P


def _cascade(prim0):
    A = _box((0, 1), (0, 1), (0, 2), 6, 6, 8)
    B = _box((0, 1), (1, 2), (0, 1), 6, 6, 4)
    C = _box((0, 1), (1, 2), (1, 2), 6, 6, 4)
    D = _box((1, 2), (0, 1), (0, 2), 6, 6, 8)
    lines = ["4", "1 7 7 9", "2 7 7 5", "3 7 7 5", "4 7 7 9",
             "one_to_one 1 6  1 5 1 7   2 5  1 5 1 7  false",
             "one_to_one 1 6  5 9 1 7   3 5  1 5 1 7  false",
             "one_to_one 2 2  1 7 1 7   3 1  1 7 1 7  false",
             "one_to_one 1 4  1 7 1 9   4 3  1 7 1 9  false"]
    full = (1, 7)
    spec = {1: ((1, (full, full)), (2, (full, full)), (3, (full, (1, 9))), (5, ((1, 9), full))),
            2: ((1, (full, full)), (3, (full, (1, 5))), (4, (full, (1, 5))), (6, ((1, 5), full))),
            3: ((2, (full, full)), (3, (full, (1, 5))), (4, (full, (1, 5))), (6, ((1, 5), full))),
            4: ((1, (full, full)), (2, (full, full)), (4, (full, (1, 9))), (5, ((1, 9), full)),
                (6, ((1, 9), full)))}
    for b, faces in spec.items():
        for f, (r1, r2) in faces:
            lines.append(f"ff {b} {f} {r1[0]} {r1[1]} {r2[0]} {r2[1]}")
    return [A, B, C, D], _nmf("\n".join(lines) + "\n"), {"ff": Dirichlet(prim0)}


def test_cascade_split_topology():
    """A's jmax face meets B on one half and C on the other, so A is cut at k = 4; A's imax
    face meets D whole, so the cut is carried into D."""
    nodes, nmf, bcs = _cascade(np.array([1.2, 40.0, -10.0, 20.0, 1e5]))
    mesh = NM.build_mesh(nodes, nmf, bc_map=bcs, finalize=False)
    assert sorted(b.name for b in mesh.blocks) == ["b1_000", "b1_001", "b2", "b3", "b4_000", "b4_001"]
    assert sum(int(np.prod(b.n)) for b in mesh.blocks) == 6 * 6 * 8 * 2 + 6 * 6 * 4 * 2
    text = NM.describe(mesh)
    assert "b1_000" in text and text.endswith("cells: 864")


def test_swap_topology():
    n = (12, 10, 8)
    h = n[0] // 2
    ix, iy, iz = (np.arange(m + 1) for m in n)
    nodes = cases.nodes_from_index(ix, iy, iz, n)
    A = nodes[:, :h + 1]
    B = np.transpose(nodes[:, h:], (0, 2, 1, 3))[:, :, :, ::-1]
    nmf = _swap_nmf(n, h)
    mesh = NM.build_mesh([A, B], nmf, bc_map={"ffx": Dirichlet(np.array([1.2, 30.0, 0.0, 0.0, 1e5]))},
                         periodic=(("py_a", 1), ("pz", 2), ("py_b", 0)), finalize=False)
    assert len(mesh.blocks) == 2
    f = mesh.blocks[0].faces[(0, 1)]
    assert f.nb == 1 and tuple(f.src_face) == (1, 0)


def _swap_nmf(n, h):
    ny, nz, nxb = n[1] + 1, n[2] + 1, n[0] - h + 1
    return _nmf(f"""\
        2
        1 {h + 1} {ny} {nz}
        2 {ny} {nxb} {nz}
        one_to_one 1 4  1 {ny} 1 {nz}   2 5  {nz} 1 1 {ny}  true
        ffx 1 3  1 {ny} 1 {nz}
        ffx 2 6  1 {nz} 1 {ny}
        py_a 1 5  1 {nz} 1 {h + 1}
        py_a 1 6  1 {nz} 1 {h + 1}
        pz   1 1  1 {h + 1} 1 {ny}
        pz   1 2  1 {h + 1} 1 {ny}
        py_b 2 3  1 {nxb} 1 {nz}
        py_b 2 4  1 {nxb} 1 {nz}
        pz   2 1  1 {ny} 1 {nxb}
        pz   2 2  1 {ny} 1 {nxb}
        """)


# ------------------------------------------------------------------------------ GPU
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    import wmles_jax_amd as wmx
    return torch, wmx


def _gather(mesh, R):
    out = {}
    for blk, r, g in zip(mesh.blocks, R, mesh.geoms):
        xc = blk.interior(g.xc.cpu().numpy()).reshape(3, -1).T
        rr = blk.interior(r.cpu().numpy()).reshape(5, -1).T
        for c, v in zip(xc, rr):
            out[tuple(np.round(c, 8))] = v
    return out


@pytest.mark.gpu
def test_swap_and_reversed_ranges_equal_single_block():
    """Block 2 transposed and flipped (Swap=TRUE, reversed range): the assembled
    two-block RHS equals the single-block RHS."""
    torch, wmx = _gpu()
    n = (12, 10, 8)
    h = n[0] // 2
    ix, iy, iz = (np.arange(m + 1) for m in n)
    nodes = cases.nodes_from_index(ix, iy, iz, n)
    ff = wmx.Dirichlet(np.array([1.2, 30.0, 0.0, 0.0, 1e5]))
    ref = wmx.Block("ref", nodes)
    m0 = wmx.Mesh([ref])
    m0.connect(ref, "jmin", ref, "jmax", shift=(0, -cases.L[1], 0))
    m0.connect(ref, "kmin", ref, "kmax", shift=(0, 0, -cases.L[2]))
    ref.set_bc("imin", ff)
    ref.set_bc("imax", ff)
    m0.finalize()
    A = nodes[:, :h + 1]
    B = np.transpose(nodes[:, h:], (0, 2, 1, 3))[:, :, :, ::-1]
    m1 = NM.build_mesh([A, B], _swap_nmf(n, h), bc_map={"ffx": ff}, periodic=(("py_a", 1), ("pz", 2), ("py_b", 0)))
    gas = wmx.Gas()
    s0, s1 = wmx.Solver(m0, gas, wmx.KEP(4), sgs="vreman"), wmx.Solver(m1, gas, wmx.KEP(4), sgs="vreman")
    R0 = _gather(m0, s0.rhs(s0.init_state(cases.prim_fn)))
    R1 = _gather(m1, s1.rhs(s1.init_state(cases.prim_fn)))
    assert R0.keys() == R1.keys()
    scale = np.abs(np.array(list(R0.values()))).max()
    err = max(np.abs(R0[k] - R1[k]).max() for k in R0) / scale
    print("swap + reversed range vs single block:", err)
    # block 2's geometry is computed from transposed / flipped nodes, so its face vectors
    # round differently (measured 1.7e-12; the reference test allows 1e-9)
    assert err < 1e-11


def _freestream_residual(wmx, mesh, sc, prim0):
    """max over interior cells of |R| / (F0 * sum of the cell's face areas / vol): the
    free-stream residual scaled by the flux magnitude (the reference's measure,
    tests/test_plot3d_nmf.py:32-50)."""
    gas = wmx.Gas()
    s = wmx.Solver(mesh, gas, sc, sgs="vreman")
    Q = s.init_state(lambda xc: np.stack([prim0[i] + 0 * xc[0] for i in range(5)]))
    R = s.rhs(Q)
    E = prim0[4] / (gas.gamma - 1) + 0.5 * prim0[0] * np.sum(prim0[1:4] ** 2)
    F0 = (E + prim0[4]) * np.linalg.norm(prim0[1:4]) + prim0[4]
    worst = 0.0
    for blk, r, g in zip(mesh.blocks, R, mesh.geoms):
        av = 0.0
        for d in range(3):
            S = blk.slab(g.S[d].cpu().numpy(), d, 0, 1)
            lo = S[(slice(None),) + tuple(slice(None, -1) if a == d else slice(None) for a in range(3))]
            hi = S[(slice(None),) + tuple(slice(1, None) if a == d else slice(None) for a in range(3))]
            av = av + (np.linalg.norm(lo, axis=0) + np.linalg.norm(hi, axis=0))
        av = av / blk.interior(g.vol.cpu().numpy())
        res = np.abs(blk.interior(r.cpu().numpy())).max(axis=0)
        worst = max(worst, float((res / (F0 * av)).max()))
    return worst


@pytest.mark.gpu
def test_cascade_split_preserves_free_stream():
    torch, wmx = _gpu()
    prim0 = np.array([1.2, 40.0, -10.0, 20.0, 1e5])
    nodes, nmf, _ = _cascade(prim0)
    mesh = NM.build_mesh(nodes, nmf, bc_map={"ff": wmx.Dirichlet(prim0)})
    for sc in (wmx.KEP(4), wmx.MUSCLHLLC(), wmx.Hybrid()):
        worst = _freestream_residual(wmx, mesh, sc, prim0)
        print(type(sc).__name__, "free-stream residual (scaled)", worst)
        assert worst < 1e-11


# ------------------------------------------------------------------ L5 airfoil (TMBWG)
L5 = os.path.join(GOLDEN, "nak_a_ff_500c_2zn_L5")


def _l5_nodes():
    """The L5 NACA airfoil grid (data file of the reference's tests) and its NMF,
    respanned to 4 spanwise cells over 0.1."""
    grid = P3.read_plot3d(L5 + ".p3dfmt")
    assert grid.formatted and grid.ndim == 3 and grid.iblank is None
    assert [X.shape[1:] for X in grid.nodes] == [(2, 43, 57), (2, 65, 29)]
    return NM.respan(grid.nodes, NM.read_nmf(L5 + ".nmf"), ncells=4, length=0.1)


def _interior_volumes(X):
    """Cell volumes of a node array (3, ni+1, nj+1, nk+1) by the oracle's geometry."""
    import oracle.mesh as OM
    return OM.compute_geometry(np.asarray(X, dtype=float))[0]


def test_l5_airfoil_grid_assembly():
    """The reference's airfoil acceptance check on the CPU (tests/test_plot3d_nmf.py:250-268):
    the wake block is cut at the trailing-edge line, the airfoil surface is b2.kmin only,
    and the sub-blocks' cells add up to the original cells, volume included."""
    nodes, nmf = _l5_nodes()
    assert [X.shape[1:] for X in nodes] == [(5, 43, 57), (5, 65, 29)]
    assert np.allclose(nodes[0][0, -1] - nodes[0][0, 0], 0.0)        # span along the node axis 0
    mesh = NM.build_mesh(nodes, nmf, bc_map={"farfield_riem": Dirichlet([1.2, 60.0, 0.0, 15.0, 1e5]),
                                             "viscous_solid": NoSlipWall()},
                         periodic=(("symmetry_y", 0),), finalize=False)
    assert sorted(b.name for b in mesh.blocks) == ["b1_000", "b1_001", "b2"]
    assert sum(int(np.prod(b.n)) for b in mesh.blocks) == 4 * (42 * 56 + 64 * 28)
    walls = [(b.name, f) for b in mesh.blocks for f, bc in b.faces.items() if isinstance(bc, NoSlipWall)]
    assert walls == [("b2", (2, 0))]
    for b in mesh.blocks:                                  # every face is assigned
        assert len(b.faces) == 6
        assert all(isinstance(b.faces[(0, s)], NM.Interface) for s in range(2))   # periodic span
    vol = sum(float(_interior_volumes(b.x_int).sum()) for b in mesh.blocks)
    vol_ref = sum(float(_interior_volumes(X).sum()) for X in nodes)
    assert abs(vol - vol_ref) < 1e-12 * vol_ref


@pytest.mark.gpu
def test_l5_airfoil_volume_and_free_stream():
    """On the device: the assembled mesh's volumes add up to the grid's, and a uniform
    stream is preserved to < 1e-11 (scaled) by KEP(4), MUSCLHLLC and Hybrid, with the
    airfoil surface as a Dirichlet face (reference tests/test_plot3d_nmf.py:263-275)."""
    torch, wmx = _gpu()
    nodes, nmf = _l5_nodes()
    prim0 = np.array([1.2, 60.0, 0.0, 15.0, 1e5])
    ff = wmx.Dirichlet(prim0)
    mesh = NM.build_mesh(nodes, nmf, bc_map={"farfield_riem": ff, "viscous_solid": wmx.NoSlipWall()},
                         periodic=(("symmetry_y", 0),))
    vol = sum(float(b.interior(g.vol).sum()) for b, g in zip(mesh.blocks, mesh.geoms))
    vol_ref = sum(float(_interior_volumes(X).sum()) for X in nodes)
    assert abs(vol - vol_ref) < 1e-12 * vol_ref
    mesh_fs = NM.build_mesh(nodes, nmf, bc_map={"farfield_riem": ff, "viscous_solid": ff},
                            periodic=(("symmetry_y", 0),))
    for sc in (wmx.KEP(4), wmx.MUSCLHLLC(), wmx.Hybrid()):
        worst = _freestream_residual(wmx, mesh_fs, sc, prim0)
        print("L5 airfoil", type(sc).__name__, "free-stream residual (scaled)", worst)
        assert worst < 1e-11

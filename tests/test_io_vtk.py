"""Checkpoint files: the VTK XML writer / parser and the hexahedral layout (CPU), the
restart round trips through a solver (GPU)."""
import base64
import os
import zlib

import numpy as np
import pytest

from wmles_jax_amd import io_vtk as IO


def _hex_volume(P):
    """Volume of a hexahedron in VTK node order by its 6-tetrahedron split."""
    tets = ((0, 1, 3, 4), (1, 2, 3, 6), (1, 3, 4, 6), (1, 4, 5, 6), (3, 4, 6, 7), (4, 6, 7, 7))
    v = 0.0
    for a, b, c, d in tets[:5]:
        v += np.dot(P[b] - P[a], np.cross(P[c] - P[a], P[d] - P[a])) / 6.0
    return v


def test_hex_connectivity_layout():
    ni, nj, nk = 3, 2, 4
    conn = IO.hex_connectivity(ni, nj, nk)
    assert conn.shape == (ni * nj * nk, 8)
    s = (nj + 1) * (nk + 1)
    assert conn[0].tolist() == [0, s, s + nk + 1, nk + 1, 1, s + 1, s + nk + 2, nk + 2]
    # cell (i, j, k) in C order: the last cell's first node is node (ni-1, nj-1, nk-1)
    assert conn[-1, 0] == ((ni - 1) * (nj + 1) + (nj - 1)) * (nk + 1) + (nk - 1)
    # right-handed unit hexahedra on a uniform grid: every volume is +1
    x = np.stack(np.meshgrid(np.arange(ni + 1.0), np.arange(nj + 1.0), np.arange(nk + 1.0), indexing="ij"))
    pts = np.moveaxis(x, 0, -1).reshape(-1, 3)
    vols = [_hex_volume(pts[c]) for c in conn]
    np.testing.assert_allclose(vols, 1.0, rtol=1e-14)


def test_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    ni, nj, nk = 2, 3, 2
    pts = rng.random(((ni + 1) * (nj + 1) * (nk + 1), 3))
    conn = IO.hex_connectivity(ni, nj, nk)
    cells = {"rho": rng.random(12), "velocity": rng.random((12, 3)), "pressure": 1e5 + rng.random(12)}
    path = os.path.join(tmp_path, "c_block0_000007.vtu")
    IO.write_vtu(path, pts, conn, cells, {"step": np.asarray([7], np.int64), "time": np.asarray([0.25])})
    g = IO.read_vtu(path)
    assert np.array_equal(g["points"], pts)
    assert np.array_equal(g["connectivity"], conn)
    assert np.all(g["types"] == IO.VTK_HEXAHEDRON)
    for k, v in cells.items():
        assert np.array_equal(g["cells"][k], v)
    assert int(g["field"]["step"][0]) == 7 and float(g["field"]["time"][0]) == 0.25


def _el(text, dtype="Float64", fmt="binary", ncomp=1):
    import xml.etree.ElementTree as ET
    el = ET.Element("DataArray", type=dtype, format=fmt, NumberOfComponents=str(ncomp))
    el.text = text
    return el


def test_parser_reads_the_vtk_encodings():
    """ASCII; uncompressed binary with the header encoded separately (UInt32); zlib
    blocks with a separately encoded UInt64 header (VTK's compressed layout)."""
    a = np.linspace(-1.0, 2.0, 11)
    assert np.array_equal(IO._read_array(_el(" ".join(repr(float(x)) for x in a), fmt="ascii"), "UInt32", False), a)
    raw = a.tobytes()
    sep = base64.b64encode(np.uint32(len(raw)).tobytes()).decode() + base64.b64encode(raw).decode()
    assert np.array_equal(IO._read_array(_el(sep), "UInt32", False), a)
    blocks = [raw[:40], raw[40:80], raw[80:]]
    comp = [zlib.compress(b) for b in blocks]
    head = np.asarray([len(blocks), 40, len(blocks[-1])] + [len(c) for c in comp], dtype=np.uint64)
    text = base64.b64encode(head.tobytes()).decode() + base64.b64encode(b"".join(comp)).decode()
    assert np.array_equal(IO._read_array(_el(text), "UInt64", True), a)
    v = np.arange(12, dtype=np.float64)
    out = IO._read_array(_el(base64.b64encode(np.uint64(96).tobytes() + v.tobytes()).decode(), ncomp=3),
                         "UInt64", False)
    assert out.shape == (4, 3) and np.array_equal(out.reshape(-1), v)


def test_pvd_entries_and_identity(tmp_path):
    p = os.path.join(tmp_path, "case.pvd")
    with open(p, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1"><Collection>'
                '<DataSet timestep="0.5" part="0" file="case_block0_000010.vtu"/>'
                '<DataSet timestep="1" part="0" file="case_block1_000020.vtu"/></Collection></VTKFile>')
    assert IO._pvd_entries(p) == [(0.5, "case_block0_000010.vtu"), (1.0, "case_block1_000020.vtu")]
    assert IO._block_file_key("/x/case_block12_000345.vtu") == ("case", 12, 345)
    assert IO._block_file_key("/x/my_case_block3_7.vtu") == ("my_case", 3, 7)
    assert IO._block_file_key("/x/case_000345.vtu") is None
    # step selection: the collection's latest step; a block file's own step and its siblings
    for b, st in ((0, 10), (1, 10), (0, 20), (1, 20)):
        open(os.path.join(tmp_path, f"case_block{b}_{st:06d}.vtu"), "w").close()
    with open(p, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1"><Collection>' +
                "".join(f'<DataSet timestep="{st / 10}" part="0" file="case_block{b}_{st:06d}.vtu"/>'
                        for st in (10, 20) for b in (0, 1)) + '</Collection></VTKFile>')
    step, files = IO._checkpoint_files(p)
    assert step == 20 and sorted(files) == [0, 1] and files[1][1] == 2.0
    step, files = IO._checkpoint_files(os.path.join(tmp_path, "case_block1_000010.vtu"))
    assert step == 10 and sorted(files) == [0, 1] and files[0][1] is None
    with pytest.raises(ValueError):
        IO._checkpoint_files(os.path.join(tmp_path, "case_000345.vtu"))
    with pytest.raises(FileNotFoundError):
        IO._checkpoint_files(os.path.join(tmp_path, "other_block0_000010.vtu"))


@pytest.mark.gpu
def test_restart_round_trips_through_a_solver(tmp_path):
    """write_case -> read_restart (PVD and block VTU) and the npz form on a two-block
    channel after two steps: the same interior state (npz: bit for bit; VTU through the
    primitive fields: to 1e-14) with the ghosts of the solver's own sweep."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    import cases
    import wmles_jax_amd as wmx
    spec, arrays = cases.channel_spec(2, 24, 16, 8, {"kind": "KEP", "order": 2}, "wale")
    mesh, s = cases.build(wmx, spec, arrays)
    pr = spec["props"]
    Q = s.init_state(lambda xc: cases.channel_prim(xc, pr, 0))
    t = 0.0
    for _ in range(2):
        dt = float(s.max_dt(Q))
        Q = s.advance(Q, dt)
        t += dt
    ref = s.apply_bcs(Q)
    w = IO.write_case(mesh, s, Q, 2, t, str(tmp_path), name="chan")
    IO.write_case(mesh, s, Q, 5, t + 1.0, str(tmp_path), name="chan", writer=w)
    for path, step, time in ((os.path.join(tmp_path, "chan.pvd"), 5, t + 1.0),
                             (os.path.join(tmp_path, "chan_block1_000002.vtu"), 2, t)):
        Qr, st, tm = IO.read_restart(path, s)
        assert st == step and abs(tm - time) <= 1e-15 * time
        for blk, a, b in zip(mesh.blocks, Qr, ref):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            for v in range(5):
                np.testing.assert_allclose(blk.interior(a[v]), blk.interior(b[v]), rtol=1e-14,
                                           atol=1e-14 * np.abs(b[v]).max())
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())
    npz = os.path.join(tmp_path, "chan.npz")
    IO.save_npz_restart(npz, s, Q, 2, t)
    Qn, st, tm = IO.load_npz_restart(npz, s)
    assert st == 2 and tm == t
    for a, b in zip(Qn, ref):
        assert torch.equal(a, b)
    # a collection whose latest step lacks a block is refused with the reference's message
    one = os.path.join(tmp_path, "one.pvd")
    with open(one, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1"><Collection>'
                f'<DataSet timestep="{t + 1.0!r}" part="0" file="chan_block0_000005.vtu"/>'
                '</Collection></VTKFile>')
    with pytest.raises(ValueError, match="holds blocks"):
        IO.read_restart(one, s)

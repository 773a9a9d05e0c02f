"""C ABI and host-side logic checks that run without a GPU."""
import os
import re

import numpy as np
import pytest

import cases
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|int32_t)\s+(wmx_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "wmles_jax_amd", "libwmles_amd.so"))
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header_sizes():
    from wmles_jax_amd import _lib
    import ctypes
    assert ctypes.sizeof(_lib.FaceDesc) == 120
    assert ctypes.sizeof(_lib.BlockDesc) == 16 + 6 * 120 + 8 * 8


def test_ctypes_structs_match_the_c_header(tmp_path):
    """sizeof / offsetof of every field of the ctypes mirrors equal the C compiler's."""
    import ctypes
    import subprocess
    from wmles_jax_amd import _lib
    structs = {"wmx_face_desc": _lib.FaceDesc, "wmx_block_desc": _lib.BlockDesc,
               "wmx_numerics_desc": _lib.NumericsDesc, "wmx_inflow_desc": _lib.InflowDesc,
               "wmx_wall_net_desc": _lib.WallNetDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "wmles_amd.h"', "int main(void) {"]
    expect = []
    for cname, py in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(py))
        for fname, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(py, fname).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == expect


@pytest.mark.parametrize("kind", ["split", "flip", "rot_cyc"])
def test_orientation_matching_equals_oracle(kind):
    """Host-side interface orientation (mesh.py:169-193) equals the oracle's."""
    import wmles_jax_amd.mesh as pm
    spec, arrays = cases.multiblock_spec(kind, (12, 10, 8))
    ob = {n: oracle.Block(n, arrays[f"nodes_{n}"]) for n in spec["blocks"]}
    pb = {n: pm.Block(n, arrays[f"nodes_{n}"]) for n in spec["blocks"]}
    om, pmesh = oracle.Mesh(list(ob.values())), pm.Mesh(list(pb.values()))
    for a, fa, b, fb, sh in spec["connect"]:
        om.connect(ob[a], fa, ob[b], fb, shift=sh)
        pmesh.connect(pb[a], fa, pb[b], fb, shift=sh)
    for n in spec["blocks"]:
        for f, itf in ob[n].faces.items():
            p = pb[n].faces[f]
            assert (p.nb, p.src_face, p.mapping, p.flip) == (itf.nb, itf.src_face, itf.mapping, itf.flip)
            np.testing.assert_array_equal(p.shift, itf.shift)


def test_mismatched_interface_raises():
    import wmles_jax_amd.mesh as pm
    spec, arrays = cases.multiblock_spec("split", (12, 10, 8))
    A, B = pm.Block("A", arrays["nodes_A"]), pm.Block("B", arrays["nodes_B"])
    m = pm.Mesh([A, B])
    with pytest.raises(ValueError):
        m.connect(A, "imax", B, "imin", shift=(0.1, 0.0, 0.0))


def test_face_descriptor_packing():
    import wmles_jax_amd as wmx
    from wmles_jax_amd.mesh import block_desc
    spec, arrays = cases.multiblock_spec("rot_cyc", (12, 10, 8), ybc={"kind": "WallModel", "T_wall": 300.0,
                                                                     "n_match": 2})
    blocks = {n: wmx.Block(n, arrays[f"nodes_{n}"]) for n in spec["blocks"]}
    m = wmx.Mesh(list(blocks.values()))
    for a, fa, b, fb, sh in spec["connect"]:
        m.connect(blocks[a], fa, blocks[b], fb, shift=sh)
    for nm, face, bc in spec["bcs"]:
        blocks[nm].set_bc(face, cases.make_bc(wmx, bc))
    d = block_desc(blocks["B"], m.index)
    itf = blocks["B"].faces[(2, 0)]
    fd = d.faces[4]
    assert fd.kind == 0 and fd.nb == itf.nb and tuple(fd.mapping) == itf.mapping
    wmf = d.faces[0]
    assert wmf.kind == 6 and wmf.T_wall == 300.0 and wmf.n_match == 2


def test_ensemble_descriptor_merge_offsets_neighbours():
    """Merged ensemble plans: case ids and interface neighbours shifted per case."""
    from wmles_jax_amd import _lib
    from wmles_jax_amd.ensemble import merge_block_descs

    def case(nblk):
        arr = (_lib.BlockDesc * nblk)()
        for b in range(nblk):
            for f in range(6):
                arr[b].faces[f].kind = _lib.BC_SLIP
            arr[b].faces[0].kind = _lib.FACE_INTERFACE        # imin <- previous block
            arr[b].faces[0].nb = (b - 1) % nblk
            arr[b].faces[1].kind = _lib.FACE_INTERFACE        # imax <- next block
            arr[b].faces[1].nb = (b + 1) % nblk
            arr[b].n[0] = 10 + b
        return arr

    descs, offs = merge_block_descs([case(2), case(3), case(1)])
    assert offs == [0, 2, 5]
    assert [d.case_id for d in descs] == [0, 0, 1, 1, 1, 2]
    assert [d.faces[0].nb for d in descs] == [1, 0, 4, 2, 3, 5]
    assert [d.faces[1].nb for d in descs] == [1, 0, 3, 4, 2, 5]
    assert [d.n[0] for d in descs] == [10, 11, 10, 11, 12, 10]
    assert all(d.faces[2].nb == 0 for d in descs)             # non-interface faces untouched


def test_tile_model_picks_the_shortest_makespan():
    """Automatic i-planes per tile: ceil(T / 2 CUs) rounds of (ci + alpha) planes."""
    import ctypes as C
    from wmles_jax_amd import _lib

    def model(dims, prec, cus=256):
        arr = (C.c_int32 * len(dims))(*dims)
        ci, nt = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib.wmx_tile_model(len(dims) // 3, arr, prec, cus, C.byref(ci), C.byref(nt)))
        return ci.value, nt.value

    def brute(dims, prec, cus=256):
        tj, alpha = 15, (1.5 if prec == _lib.PREC_FP32_BULK else 0.5)
        best = None
        for ci in range(64, 3, -1):
            T = sum(-(-dims[3 * b] // ci) * -(-dims[3 * b + 1] // tj) * -(-dims[3 * b + 2] // 16)
                    for b in range(len(dims) // 3))
            cost = -(-T // (2 * cus)) * (ci + alpha)
            if best is None or cost < best[0] - 1e-9:
                best = (cost, ci, T)
        return best[1], best[2]

    for dims in ([256, 128, 128], [512, 256, 256], [512, 128, 64], [72, 56, 40, 33, 39, 56],
                 [100, 30, 17]):
        for prec in (_lib.PREC_FP64, _lib.PREC_FP32_BULK):
            assert model(dims, prec) == brute(dims, prec)
    # the 256 x 128 x 128 channel: 72 tiles of 15 x 16 per i chunk; 7 chunks of 37 planes
    # make 504 tiles, one round of 512 slots (19 planes: 1008 tiles, two rounds, 41 plane
    # times against 38.5)
    assert model([256, 128, 128], _lib.PREC_FP32_BULK) == (37, 504)
    # the 512 x 256 x 256 channel keeps 32 planes: 4608 tiles, nine rounds
    assert model([512, 256, 256], _lib.PREC_FP32_BULK) == (32, 4608)

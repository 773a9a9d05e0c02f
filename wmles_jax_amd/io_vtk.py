"""Checkpoint / restart: VTU + PVD in the layout of the reference's `wmles/io_vtk.py`, and
a raw interior `.npz` form.  No pyvista: the VTK XML files are written and parsed here.

Files (reference `io_vtk.py:60-193,296-329`):
* one `<case>_block<index>_<step:06d>.vtu` per block: an UnstructuredGrid of the block's
  physical cells as VTK hexahedra over its interior nodes (points in C order of
  (i, j, k) nodes, k fastest; hexahedron node order (0,0,0) (1,0,0) (1,1,0) (0,1,0)
  (0,0,1) (1,0,1) (1,1,1) (0,1,1)), cell data `rho`, `velocity` (3 components),
  `pressure`, `temperature`, `mach`, field data `step` and `time`;
* `<case>.pvd`: the time collection, one DataSet per block file.

`read_restart(path, solver)` accepts a `.pvd` (latest step) or any block `.vtu` (its
sibling blocks of the same step) and returns `(Q, step, time)` with the ghost cells
filled by the solver's boundary sweep, like the reference.  It checks block indices,
step numbers, cell counts, points (rtol 1e-10) and connectivity, the fields and their
validity, and time consistency, with the reference's error conditions.  The parser
reads ASCII and inline binary data arrays, uncompressed or zlib-compressed (VTK's block
format), with UInt32 or UInt64 headers -- the forms VTK writers produce.

The `.npz` form (`save_npz_restart` / `load_npz_restart`) stores each block's interior
conserved state (5, n_i, n_j, n_k) in float64 with `step`, `time` and the block shapes.
"""
import base64
import glob
import os
import re
import xml.etree.ElementTree as ET
import zlib
from pathlib import Path

import numpy as np
import torch

VTK_HEXAHEDRON = 12
_NP = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32,
       "UInt8": np.uint8, "UInt32": np.uint32, "UInt64": np.uint64, "Int8": np.int8,
       "UInt16": np.uint16, "Int16": np.int16}
_VTK = {np.dtype(v).str: k for k, v in _NP.items()}


# ------------------------------------------------------------------------------ geometry
def hex_connectivity(ni, nj, nk):
    """(ncell, 8) node ids of the hexahedra of an (ni, nj, nk) block, cells in C order."""
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    corners = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
    ids = [(((i + a) * (nj + 1) + (j + b)) * (nk + 1) + (k + c)).ravel() for a, b, c in corners]
    return np.stack(ids, axis=1).astype(np.int64)


def block_points(blk):
    return np.ascontiguousarray(np.moveaxis(np.asarray(blk.x_int, dtype=np.float64), 0, -1).reshape(-1, 3))


# ------------------------------------------------------------------------------ XML I/O
def _b64_array(a):
    a = np.ascontiguousarray(a)
    raw = a.tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _data_array(parent, name, a, ncomp=1):
    a = np.ascontiguousarray(a)
    el = ET.SubElement(parent, "DataArray", type=_VTK[a.dtype.str], format="binary")
    if name:
        el.set("Name", name)
    if ncomp > 1:
        el.set("NumberOfComponents", str(ncomp))
    el.text = _b64_array(a)
    return el


def write_vtu(path, points, conn, cell_fields, field_data):
    """Write an UnstructuredGrid of hexahedra; cell_fields {name: (ncell,) or (ncell, 3)}."""
    ncell = len(conn)
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    fd = ET.SubElement(grid, "FieldData")
    for name, arr in field_data.items():
        el = _data_array(fd, name, np.asarray(arr))
        el.set("NumberOfTuples", str(np.asarray(arr).size))
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(len(points)), NumberOfCells=str(ncell))
    pts = ET.SubElement(piece, "Points")
    _data_array(pts, "Points", points.astype(np.float64), 3)
    cells = ET.SubElement(piece, "Cells")
    _data_array(cells, "connectivity", conn.reshape(-1).astype(np.int64))
    _data_array(cells, "offsets", (8 * np.arange(1, ncell + 1)).astype(np.int64))
    _data_array(cells, "types", np.full(ncell, VTK_HEXAHEDRON, dtype=np.uint8))
    cd = ET.SubElement(piece, "CellData")
    for name, arr in cell_fields.items():
        arr = np.asarray(arr)
        _data_array(cd, name, arr, 1 if arr.ndim == 1 else arr.shape[1])
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def _decode_binary(text, dtype, header, compressed):
    text = "".join(text.split())
    hb = 8 if header == "UInt64" else 4
    hdt = np.uint64 if hb == 8 else np.uint32
    if not compressed:
        head = base64.b64decode(text[:4 * ((hb + 2) // 3)])[:hb]
        nbytes = int(np.frombuffer(head, dtype=hdt)[0])
        whole = base64.b64decode(text)
        if len(whole) >= hb + nbytes and len(text) == 4 * ((hb + nbytes + 2) // 3):
            raw = whole[hb:hb + nbytes]                     # header and data encoded together
        else:                                               # encoded one after the other
            hchars = 4 * ((hb + 2) // 3)
            raw = base64.b64decode(text[hchars:])[:nbytes]
        return np.frombuffer(raw, dtype=dtype)
    nb = int(np.frombuffer(base64.b64decode(text[:4 * ((hb + 2) // 3)])[:hb], dtype=hdt)[0])
    hlen = hb * (3 + nb)
    hchars = 4 * ((hlen + 2) // 3)
    h = np.frombuffer(base64.b64decode(text[:hchars])[:hlen], dtype=hdt).astype(np.int64)
    sizes = h[3:]
    data = base64.b64decode(text[hchars:])
    out, o = [], 0
    for sz in sizes:
        out.append(zlib.decompress(data[o:o + sz]))
        o += sz
    return np.frombuffer(b"".join(out), dtype=dtype)


def _read_array(el, header, compressed):
    dtype = np.dtype(_NP[el.get("type")]).newbyteorder("<")
    fmt = el.get("format", "ascii")
    if fmt == "ascii":
        a = np.array((el.text or "").split(), dtype=dtype)
    elif fmt == "binary":
        a = _decode_binary(el.text or "", dtype, header, compressed)
    else:
        raise ValueError(f"unsupported DataArray format '{fmt}' (appended data is not read)")
    nc = int(el.get("NumberOfComponents", "1"))
    return a.reshape(-1, nc) if nc > 1 else a


def read_vtu(path):
    """{points, connectivity (ncell, 8), types, cells: {name: array}, field: {name: array}}."""
    try:
        root = ET.parse(path).getroot()
    except (ET.ParseError, OSError) as exc:
        raise ValueError(f"could not read VTU {path}: {exc}") from exc
    if root.get("type") != "UnstructuredGrid":
        raise ValueError(f"{path}: not a VTK UnstructuredGrid file")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError(f"{path}: big-endian VTU is not supported")
    header = root.get("header_type", "UInt32")
    compressed = root.get("compressor") is not None
    grid = root.find("UnstructuredGrid")
    piece = grid.find("Piece")
    rd = lambda el: _read_array(el, header, compressed)
    out = {"field": {}, "cells": {}}
    fd = grid.find("FieldData")
    if fd is not None:
        for el in fd.findall("DataArray"):
            out["field"][el.get("Name")] = rd(el)
    out["points"] = rd(piece.find("Points").find("DataArray")).reshape(-1, 3)
    ce = {el.get("Name"): rd(el) for el in piece.find("Cells").findall("DataArray")}
    off = ce["offsets"].astype(np.int64)
    conn = ce["connectivity"].astype(np.int64)
    if len(off) and np.all(np.diff(np.concatenate([[0], off])) == 8):
        out["connectivity"] = conn.reshape(-1, 8)
    else:
        out["connectivity"] = None
    out["types"] = ce["types"].astype(np.int64)
    cd = piece.find("CellData")
    if cd is not None:
        for el in cd.findall("DataArray"):
            out["cells"][el.get("Name")] = rd(el)
    return out


# ------------------------------------------------------------------------------ fields
def restart_fields(solver, Q):
    """{block index: {name: ghost-free array}}: rho, velocity (3, n), pressure, temperature,
    mach, from the FP64 state."""
    gas = solver.gas
    out = {}
    for ib, (blk, q) in enumerate(zip(solver.blocks, solver._as_state(Q))):
        w, T = gas.cons_to_prim(q)
        c = gas.sound_speed(T)
        M = torch.sqrt((w[1] * w[1] + w[2] * w[2]) + w[3] * w[3]) / c
        cpu = lambda a: np.asarray(blk.interior(a).detach().cpu())
        out[ib] = {"rho": cpu(w[0]), "velocity": cpu(w[1:4]), "pressure": cpu(w[4]),
                   "temperature": cpu(T), "mach": cpu(M)}
    return out


def _cell_major(arr, n):
    arr = np.asarray(arr)
    if arr.shape == tuple(n):
        return arr.reshape(-1)
    return np.stack([c.reshape(-1) for c in arr], axis=1)


# ------------------------------------------------------------------------------ writer
def _pvd_entries(path):
    try:
        root = ET.parse(path).getroot()
        entries = [(float(d.get("timestep")), d.get("file")) for d in root.iter("DataSet")
                   if d.get("file") is not None and d.get("timestep") is not None]
    except (ET.ParseError, OSError, ValueError) as exc:
        raise ValueError(f"could not read VTU collection {path}: {exc}") from exc
    if not entries:
        raise ValueError(f"{path}: no time-step datasets found")
    return entries


class VTKWriter:
    """Holds the .pvd collection of one case across time steps (io_vtk.py:296-329)."""

    def __init__(self, outdir, name="case", resume=False):
        self.outdir, self.name = str(outdir), name
        os.makedirs(self.outdir, exist_ok=True)
        pvd = os.path.join(self.outdir, f"{name}.pvd")
        self.entries = _pvd_entries(pvd) if resume else []

    def _write_pvd(self):
        lines = ['<?xml version="1.0"?>',
                 '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">', "  <Collection>"]
        lines += [f'    <DataSet timestep="{t:.9g}" part="0" file="{p}"/>' for t, p in self.entries]
        lines += ["  </Collection>", "</VTKFile>", ""]
        with open(os.path.join(self.outdir, f"{self.name}.pvd"), "w") as f:
            f.write("\n".join(lines))

    def write_step(self, mesh, solver, Q, step, t, extra_fields=None):
        fields = restart_fields(solver, Q)
        for ib, fd in (extra_fields or {}).items():
            fields.setdefault(ib, {}).update(fd)
        for ib, blk in enumerate(mesh.blocks):
            rel = f"{self.name}_block{ib}_{int(step):06d}.vtu"
            cells = {k: _cell_major(v, blk.n) for k, v in fields[ib].items()}
            write_vtu(os.path.join(self.outdir, rel), block_points(blk), hex_connectivity(*blk.n), cells,
                      {"step": np.asarray([int(step)], dtype=np.int64),
                       "time": np.asarray([float(t)], dtype=np.float64)})
            self.entries.append((float(t), rel))
        self._write_pvd()
        return self


def write_case(mesh, solver, Q, step, t, outdir, name="case", writer=None):
    """One .vtu per block for this step, appended to `<outdir>/<name>.pvd`."""
    writer = writer or VTKWriter(outdir, name)
    return writer.write_step(mesh, solver, Q, step, t)


# ------------------------------------------------------------------------------ reader
_NAME = re.compile(r"^(?P<case>.+)_block(?P<block>[0-9]+)_(?P<step>[0-9]+)\.vtu$")


def _block_file_key(path):
    """(case, block index, step) from a block file name `<case>_block<i>_<step>.vtu`, or
    None when the name does not follow it."""
    m = _NAME.match(Path(path).name)
    return None if m is None else (m["case"], int(m["block"]), int(m["step"]))


def _checkpoint_files(path):
    """The block files of one saved step: (step, {block: (path, time from the collection
    or None)}).  A `.pvd` gives its latest step; a block `.vtu` gives its own step, with
    the sibling files of the same case and step in its directory."""
    path = Path(path)
    if path.suffix.lower() == ".pvd":
        by_step = {}
        for t, rel in _pvd_entries(path):
            fp = (path.parent / rel).resolve()
            key = _block_file_key(fp)
            if key is not None:
                by_step.setdefault(key[2], []).append((key[1], fp, t))
        if not by_step:
            raise ValueError(f"{path}: none of the collection's datasets is a block file "
                             "'<case>_block<i>_<step>.vtu'")
        step = max(by_step)
        entries = by_step[step]
    else:
        key = _block_file_key(path)
        if key is None:
            raise ValueError(f"{path}: a restart file is a '.pvd' collection or a block file named "
                             "'<case>_block<i>_<step>.vtu'")
        case, _, step = key
        entries = []
        for fp in sorted(path.parent.glob(f"{glob.escape(case)}_block*.vtu")):
            k = _block_file_key(fp)
            if k is not None and k[0] == case and k[2] == step:
                entries.append((k[1], fp, None))
        if not entries:
            raise FileNotFoundError(f"{path}: no block files of step {step} found")
    files = {}
    for blk, fp, t in entries:
        if blk in files:
            raise ValueError(f"{path}: block {blk} appears twice in step {step}")
        files[blk] = (fp, t)
    return step, files


def _checked_primitives(fn, g, blk):
    """(5, n_i, n_j, n_k) interior primitives [rho, u, v, w, p] of one block file, after
    checking that the file is this block's grid and holds valid fields."""
    ncell = int(np.prod(blk.n))
    if len(g["types"]) != ncell:
        raise ValueError(f"{fn}: {len(g['types'])} cells, block {blk.name} has {ncell}")
    pts = block_points(blk)
    if g["points"].shape != pts.shape or not np.allclose(g["points"], pts, rtol=1e-10, atol=1e-12):
        raise ValueError(f"{fn}: the points are not block {blk.name}'s nodes")
    conn = g["connectivity"]
    if conn is None or not np.array_equal(conn, hex_connectivity(*blk.n)) or np.any(g["types"] != VTK_HEXAHEDRON):
        raise ValueError(f"{fn}: the cells are not block {blk.name}'s hexahedra")
    cells = g["cells"]
    need = [k for k in ("rho", "velocity", "pressure") if k not in cells]
    if need:
        raise ValueError(f"{fn}: cell fields {need} are missing")
    vel = np.asarray(cells["velocity"], dtype=np.float64)
    if vel.shape != (ncell, 3):
        raise ValueError(f"{fn}: velocity has shape {vel.shape}, ({ncell}, 3) expected")
    prim = np.empty((5,) + tuple(blk.n))
    prim[0] = np.asarray(cells["rho"], dtype=np.float64).reshape(blk.n)
    prim[1:4] = vel.T.reshape((3,) + tuple(blk.n))
    prim[4] = np.asarray(cells["pressure"], dtype=np.float64).reshape(blk.n)
    if not np.isfinite(prim).all() or (prim[0] <= 0).any() or (prim[4] <= 0).any():
        raise ValueError(f"{fn}: non-finite values, or density / pressure not positive")
    return prim


def read_restart(path, solver):
    """(Q, step, time) of the latest saved state of a `.pvd` collection, or of a block
    `.vtu` and its sibling blocks, with the ghosts filled by the solver's boundary sweep
    (reference `wmles/io_vtk.py:118-193`)."""
    step, files = _checkpoint_files(path)
    nb = len(solver.blocks)
    if sorted(files) != list(range(nb)):
        raise ValueError(f"{path}: step {step} holds blocks {sorted(files)}, the solver has blocks 0..{nb - 1}")
    states, times = [], []
    for ib, blk in enumerate(solver.blocks):
        fn, coll_time = files[ib]
        g = read_vtu(fn)
        states.append(_padded_state(solver, blk, _checked_primitives(fn, g, blk)))
        fstep = np.asarray(g["field"].get("step", [])).ravel()
        ftime = np.asarray(g["field"].get("time", [])).ravel()
        if fstep.size and int(fstep[0]) != step:
            raise ValueError(f"{fn}: field data says step {int(fstep[0])}, the file name {step}")
        if ftime.size:
            times.append(float(ftime[0]))
        elif coll_time is not None:
            times.append(float(coll_time))
        else:
            raise ValueError(f"{fn}: no time in the file's field data; read it through its .pvd collection")
    if not np.allclose(times, times[0], rtol=1e-9, atol=1e-15):
        raise ValueError(f"{path}: the blocks of step {step} were saved at different times {times}")
    return solver.apply_bcs(tuple(states)), step, times[0]


def _padded_state(solver, blk, prim):
    """Ghost-padded conserved state (zeros in the ghosts) from interior primitives."""
    w = torch.as_tensor(prim, dtype=torch.float64, device=solver._dev)
    q = solver.gas.prim_to_cons(w)
    ng = blk.ng
    return torch.nn.functional.pad(q, (ng, ng, ng, ng, ng, ng))


# ------------------------------------------------------------------------------ npz
def save_npz_restart(path, solver, Q, step, time):
    """Interior conserved states (float64) + step + time in one .npz."""
    Q = solver._as_state(Q)
    arrs = {f"q{ib}": np.asarray(blk.interior(q).detach().cpu()) for ib, (blk, q) in enumerate(zip(solver.blocks, Q))}
    np.savez(path, step=np.int64(step), time=np.float64(time),
             shapes=np.asarray([list(b.n) for b in solver.blocks], dtype=np.int64), **arrs)


def load_npz_restart(path, solver):
    """(Q, step, time) from save_npz_restart, ghosts filled by the boundary sweep."""
    with np.load(path, allow_pickle=False) as z:
        shapes = z["shapes"]
        if len(shapes) != len(solver.blocks) or any(tuple(s) != tuple(b.n) for s, b in zip(shapes, solver.blocks)):
            raise ValueError(f"{path}: block shapes {shapes.tolist()} do not match the solver's")
        states = []
        for ib, blk in enumerate(solver.blocks):
            q = torch.as_tensor(z[f"q{ib}"], dtype=torch.float64, device=solver._dev)
            ng = blk.ng
            states.append(torch.nn.functional.pad(q, (ng, ng, ng, ng, ng, ng)))
        step, time = int(z["step"]), float(z["time"])
    return solver.apply_bcs(tuple(states)), step, time

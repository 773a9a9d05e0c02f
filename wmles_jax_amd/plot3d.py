"""Plot3D multiblock grid files (whole format, structured), host side.

`read_plot3d(path)` -> `Plot3DGrid(nodes=[(3, ni, nj, nk) float64, ...], iblank=None or
[(ni, nj, nk) int64, ...], formatted, ndim)`; `write_plot3d(path, nodes, ...)`.  The
format follows the reference reader (`wmles/plot3d.py:1-195`):
* formatted (ASCII) or Fortran-unformatted sequential records, told apart from the
  first bytes;
* unformatted: the byte order comes from the record markers, the real size (4 or 8
  bytes) and an optional int32 iblank array from each block record's length;
* formatted: 2-D vs 3-D and iblank from the value count (Fortran `D` exponents read);
* a 2-D file (x, y per block) comes back with a singleton k axis and z = 0;
* coordinates are stored block by block, x then y (then z), i fastest.
"""
from dataclasses import dataclass
import struct

import numpy as np


@dataclass
class Plot3DGrid:
    nodes: list
    iblank: object
    formatted: bool
    ndim: int = 3


def _looks_ascii(raw):
    head = raw[:200]
    try:
        txt = head.decode("ascii")
    except UnicodeDecodeError:
        return False
    return all(ch.isprintable() or ch in "\r\n\t" for ch in txt)


def _dims(ints, nb, ndim):
    return [tuple(ints[ndim * b:ndim * (b + 1)]) + (1,) * (3 - ndim) for b in range(nb)]


def _unpack_blocks(vals, dims, ndim, with_iblank):
    nodes, ibl, o = [], [], 0
    for d in dims:
        n = int(np.prod(d))
        X = np.zeros((3,) + tuple(d))
        for c in range(ndim):
            X[c] = vals[o:o + n].reshape(d, order="F")
            o += n
        nodes.append(X)
        if with_iblank:
            ibl.append(np.asarray(vals[o:o + n]).astype(np.int64).reshape(d, order="F"))
            o += n
    if o != len(vals):
        raise ValueError("trailing values in Plot3D file")
    return nodes, (ibl if with_iblank else None)


def _read_ascii(text, ndim):
    tok = text.replace("D", "E").replace("d", "e").split()
    nb = int(tok[0])
    if ndim is None:
        for nd in (3, 2):
            try:
                dd = _dims([int(t) for t in tok[1:1 + nb * nd]], nb, nd)
            except ValueError:
                continue
            npts = sum(int(np.prod(d)) for d in dd)
            if len(tok) - 1 - nb * nd in (nd * npts, (nd + 1) * npts):
                ndim = nd
                break
        else:
            raise ValueError("cannot decide whether this Plot3D file is 2-D or 3-D")
    dims = _dims([int(t) for t in tok[1:1 + nb * ndim]], nb, ndim)
    npts = sum(int(np.prod(d)) for d in dims)
    vals = np.array(tok[1 + nb * ndim:], dtype=np.float64)
    if len(vals) == ndim * npts:
        ib = False
    elif len(vals) == (ndim + 1) * npts:
        ib = True
    else:
        raise ValueError(f"value count {len(vals)} matches neither {ndim * npts} (no iblank) "
                         f"nor {(ndim + 1) * npts} (iblank)")
    nodes, ibl = _unpack_blocks(vals, dims, ndim, ib)
    return Plot3DGrid(nodes, ibl, True, ndim)


def _records(raw):
    """Fortran sequential records; the byte order is the one whose markers frame the file."""
    for bo in ("<", ">"):
        recs, pos = [], 0
        while pos + 4 <= len(raw):
            n = struct.unpack(bo + "i", raw[pos:pos + 4])[0]
            if n < 0 or pos + 8 + n > len(raw) or struct.unpack(bo + "i", raw[pos + 4 + n:pos + 8 + n])[0] != n:
                break
            recs.append(raw[pos + 4:pos + 4 + n])
            pos += 8 + n
        if recs and pos == len(raw):
            return recs, bo
    raise ValueError("not a valid Fortran-unformatted Plot3D file")


def _read_binary(raw, ndim):
    recs, bo = _records(raw)
    nb = struct.unpack(bo + "i", recs[0])[0]
    nint = len(recs[1]) // 4
    if ndim is None:
        if nint == 3 * nb:
            ndim = 3
        elif nint == 2 * nb:
            ndim = 2
        else:
            raise ValueError("dimension record length matches neither 2-D nor 3-D")
    dims = _dims(list(struct.unpack(bo + f"{nint}i", recs[1])), nb, ndim)
    body = recs[2:]
    if len(body) != nb:
        raise ValueError(f"expected {nb} coordinate records, found {len(body)} "
                         "(only whole-format multiblock files are supported)")
    nodes, ibl, any_ibl = [], [], False
    for d, rec in zip(dims, body):
        n = int(np.prod(d))
        for size, code in ((4, "f4"), (8, "f8")):
            if len(rec) == ndim * n * size:
                vals, has = np.frombuffer(rec, dtype=bo + code).astype(np.float64), False
                break
            if len(rec) == ndim * n * size + 4 * n:
                cut = ndim * n * size
                vals = np.concatenate([np.frombuffer(rec[:cut], dtype=bo + code).astype(np.float64),
                                       np.frombuffer(rec[cut:], dtype=bo + "i4").astype(np.float64)])
                has = True
                break
        else:
            raise ValueError(f"record of {len(rec)} bytes doesn't fit dims {d}")
        X, ib = _unpack_blocks(vals, [d], ndim, has)
        nodes += X
        if has:
            ibl += ib
            any_ibl = True
    return Plot3DGrid(nodes, ibl if any_ibl else None, False, ndim)


def read_plot3d(path, formatted=None, ndim=None):
    """Whole-format multiblock Plot3D grid; `formatted` / `ndim` detected when None."""
    with open(path, "rb") as f:
        raw = f.read()
    if formatted is None:
        formatted = _looks_ascii(raw)
    return _read_ascii(raw.decode("ascii"), ndim) if formatted else _read_binary(raw, ndim)


def write_plot3d(path, nodes, formatted=True, iblank=None, dtype="f8", endian="<"):
    """(3, ni, nj, nk) node arrays as a whole-format 3-D multiblock Plot3D file."""
    nodes = [np.asarray(X, dtype=np.float64) for X in nodes]
    if formatted:
        with open(path, "w") as f:
            f.write(f"{len(nodes):12d}\n")
            for X in nodes:
                f.write("".join(f"{s:12d}" for s in X.shape[1:]) + "\n")
            for b, X in enumerate(nodes):
                flat = np.concatenate([X[c].ravel(order="F") for c in range(3)])
                if iblank is not None:
                    flat = np.concatenate([flat, np.asarray(iblank[b]).ravel(order="F")])
                for k in range(0, len(flat), 3):
                    f.write("".join(f"{v:26.16E}" for v in flat[k:k + 3]) + "\n")
        return

    def record(payload):
        m = struct.pack(endian + "i", len(payload))
        return m + payload + m

    with open(path, "wb") as f:
        f.write(record(struct.pack(endian + "i", len(nodes))))
        dims = [s for X in nodes for s in X.shape[1:]]
        f.write(record(struct.pack(endian + f"{len(dims)}i", *dims)))
        for b, X in enumerate(nodes):
            body = b"".join(X[c].ravel(order="F").astype(endian + dtype).tobytes() for c in range(3))
            if iblank is not None:
                body += np.asarray(iblank[b]).ravel(order="F").astype(endian + "i4").tobytes()
            f.write(record(body))

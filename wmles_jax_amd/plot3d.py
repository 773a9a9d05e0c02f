"""Plot3D grid files: whole-format multiblock, structured (host side).

    grid = read_plot3d(path)        # Plot3DGrid(nodes, iblank, formatted, ndim)
    write_plot3d(path, nodes, formatted=True, iblank=None, dtype="f8", endian="<")

The public names, arguments and the `Plot3DGrid` fields are those of the reference
(`wmles/plot3d.py:24-29,156-168`).  What the files hold is fixed by the format:

    nblocks
    (ni, nj[, nk]) for every block
    per block: x, y[, z] over all points (i fastest), then optionally an int iblank array

either as whitespace-separated text (Fortran `D` exponents allowed) or as Fortran
unformatted sequential records (4-byte length markers around every record; one record for
nblocks, one for all dimensions, one per block).  A 2-D file returns its blocks with a
singleton k axis and z = 0.

How this reader decides what a file is:
* binary vs text: a Fortran record file starts with the 4-byte marker of the nblocks
  record, whose value is 4 in one of the two byte orders; text never starts that way;
* byte order: the one in which that first marker reads 4;
* 2-D vs 3-D: from the length of the dimension record (binary), or from which choice makes
  the value count add up (text);
* real size and iblank: from each block record's byte length, which has exactly one of
  four possible values (4- or 8-byte reals, with or without a 4-byte iblank per point).
"""
from dataclasses import dataclass
import struct

import numpy as np


@dataclass
class Plot3DGrid:
    nodes: list          # per block (3, ni, nj, nk) float64
    iblank: object       # None, or per block (ni, nj, nk) int64
    formatted: bool
    ndim: int = 3


# ------------------------------------------------------------------ binary (Fortran records)
class _RecordStream:
    """Cursor over the Fortran sequential records of a byte string."""

    def __init__(self, raw):
        self.raw = raw
        self.pos = 0
        self.order = None
        for order in ("<", ">"):
            if len(raw) >= 4 and struct.unpack_from(order + "i", raw, 0)[0] == 4:
                self.order = order
                break
        if self.order is None:
            raise ValueError("Plot3D: the file does not start with a Fortran record of one integer")

    def next(self):
        raw, o = self.raw, self.order
        if self.pos + 4 > len(raw):
            raise ValueError("Plot3D: unexpected end of file (record marker)")
        n = struct.unpack_from(o + "i", raw, self.pos)[0]
        end = self.pos + 4 + n
        if n < 0 or end + 4 > len(raw):
            raise ValueError(f"Plot3D: record at byte {self.pos} claims {n} bytes, past the end of the file")
        if struct.unpack_from(o + "i", raw, end)[0] != n:
            raise ValueError(f"Plot3D: record at byte {self.pos} has mismatched length markers")
        body = raw[self.pos + 4:end]
        self.pos = end + 4
        return body

    def at_end(self):
        return self.pos == len(self.raw)


def _fortran_block(values, shape):
    """Values stored with i fastest -> array of the given (ni, nj, nk) shape."""
    return np.asarray(values).reshape(shape[::-1]).transpose(2, 1, 0)


def _parse_binary(raw, ndim):
    rs = _RecordStream(raw)
    (nb,) = struct.unpack(rs.order + "i", rs.next())
    if nb < 1:
        raise ValueError(f"Plot3D: block count {nb}")
    drec = rs.next()
    nint = len(drec) // 4
    if ndim is None:
        if nint % nb or nint // nb not in (2, 3):
            raise ValueError(f"Plot3D: a dimension record of {nint} integers fits neither 2 nor 3 "
                             f"per block for {nb} blocks")
        ndim = nint // nb
    if nint != ndim * nb:
        raise ValueError(f"Plot3D: dimension record holds {nint} integers, {ndim * nb} expected")
    ints = struct.unpack(rs.order + f"{nint}i", drec)
    shapes = [tuple(ints[ndim * b:ndim * b + ndim]) + (1,) * (3 - ndim) for b in range(nb)]
    nodes, iblank = [], []
    for b, shape in enumerate(shapes):
        if rs.at_end():
            raise ValueError(f"Plot3D: the file ends before the record of block {b + 1} of {nb}")
        body = rs.next()
        npts = int(np.prod(shape))
        layouts = {ndim * npts * rs_ + (4 * npts if ib else 0): (rs_, ib)
                   for rs_ in (8, 4) for ib in (False, True)}
        if len(body) not in layouts:
            raise ValueError(f"Plot3D: block {b + 1} record has {len(body)} bytes; points "
                             f"{shape} allow only {sorted(layouts)}")
        real, has_ib = layouts[len(body)]
        coords = np.frombuffer(body, dtype=f"{rs.order}f{real}", count=ndim * npts)
        X = np.zeros((3,) + shape)
        for c in range(ndim):
            X[c] = _fortran_block(coords[c * npts:(c + 1) * npts], shape)
        nodes.append(X)
        if has_ib:
            ib = np.frombuffer(body, dtype=f"{rs.order}i4", count=npts, offset=ndim * npts * real)
            iblank.append(_fortran_block(ib.astype(np.int64), shape))
        else:
            iblank.append(None)
    if not rs.at_end():
        raise ValueError(f"Plot3D: data after the {nb} block records (only whole-format multiblock "
                         "files are read)")
    if any(ib is not None for ib in iblank) and any(ib is None for ib in iblank):
        raise ValueError("Plot3D: iblank present for some blocks only")
    return Plot3DGrid(nodes, iblank if iblank[0] is not None else None, False, ndim)


# ------------------------------------------------------------------ text
def _parse_text(text, ndim):
    tokens = text.translate(str.maketrans("Dd", "Ee")).split()
    if not tokens:
        raise ValueError("Plot3D: empty file")
    nb = int(tokens[0])
    if nb < 1:
        raise ValueError(f"Plot3D: block count {nb}")

    def layout(nd):
        """(shapes, iblank?) if reading the header with nd dims per block accounts for every
        token, else None."""
        try:
            ints = [int(t) for t in tokens[1:1 + nd * nb]]
        except ValueError:
            return None
        if len(ints) != nd * nb or min(ints) < 1:
            return None
        shapes = [tuple(ints[nd * b:nd * b + nd]) + (1,) * (3 - nd) for b in range(nb)]
        npts = sum(int(np.prod(s)) for s in shapes)
        rest = len(tokens) - 1 - nd * nb
        if rest == nd * npts:
            return shapes, False
        if rest == (nd + 1) * npts:
            return shapes, True
        return None

    cands = [(nd, layout(nd)) for nd in ((ndim,) if ndim else (3, 2))]
    cands = [(nd, lay) for nd, lay in cands if lay is not None]
    if not cands:
        raise ValueError("Plot3D: the value count fits no 2-D or 3-D reading of the header"
                         if ndim is None else f"Plot3D: the value count does not fit a {ndim}-D header")
    nd, (shapes, has_ib) = cands[0]
    vals = np.array(tokens[1 + nd * nb:], dtype=np.float64)
    nodes, iblank, at = [], [], 0
    for shape in shapes:
        npts = int(np.prod(shape))
        X = np.zeros((3,) + shape)
        for c in range(nd):
            X[c] = _fortran_block(vals[at:at + npts], shape)
            at += npts
        nodes.append(X)
        if has_ib:
            iblank.append(_fortran_block(vals[at:at + npts].astype(np.int64), shape))
            at += npts
    return Plot3DGrid(nodes, iblank if has_ib else None, True, nd)


def _is_record_file(raw):
    return len(raw) >= 4 and 4 in (struct.unpack_from("<i", raw, 0)[0], struct.unpack_from(">i", raw, 0)[0])


def read_plot3d(path, formatted=None, ndim=None):
    """Whole-format multiblock Plot3D grid.  `formatted` (text) and `ndim` (2 or 3) are
    detected when None (reference `wmles/plot3d.py:156-165`)."""
    with open(path, "rb") as f:
        raw = f.read()
    if formatted is None:
        formatted = not _is_record_file(raw)
    if formatted:
        return _parse_text(raw.decode("ascii"), ndim)
    return _parse_binary(raw, ndim)


# ------------------------------------------------------------------ writer
def write_plot3d(path, nodes, formatted=True, iblank=None, dtype="f8", endian="<"):
    """(3, ni, nj, nk) node arrays as a whole-format 3-D multiblock Plot3D file: text
    (17 significant digits, four values per line) or Fortran unformatted records of
    `dtype` reals in byte order `endian` (reference `wmles/plot3d.py:168-195`)."""
    nodes = [np.asarray(X, dtype=np.float64) for X in nodes]
    ib = None if iblank is None else [np.asarray(a) for a in iblank]

    def block_values(b):
        """x, y, z, then iblank, each with i fastest (Fortran order)."""
        X = nodes[b]
        parts = [X[c].transpose(2, 1, 0).reshape(-1) for c in range(3)]
        return parts, (ib[b].transpose(2, 1, 0).reshape(-1) if ib is not None else None)

    if formatted:
        out = [str(len(nodes))]
        out += [" ".join(str(int(s)) for s in X.shape[1:]) for X in nodes]
        for b in range(len(nodes)):
            parts, ibv = block_values(b)
            vals = np.concatenate(parts)
            for s in range(0, len(vals), 4):
                out.append(" ".join(f"{v:.16e}" for v in vals[s:s + 4]))
            if ibv is not None:
                for s in range(0, len(ibv), 16):
                    out.append(" ".join(str(int(v)) for v in ibv[s:s + 16]))
        with open(path, "w") as f:
            f.write("\n".join(out) + "\n")
        return

    def framed(payload):
        marker = struct.pack(endian + "i", len(payload))
        return b"".join((marker, payload, marker))

    chunks = [framed(struct.pack(endian + "i", len(nodes))),
              framed(np.asarray([s for X in nodes for s in X.shape[1:]], dtype=endian + "i4").tobytes())]
    for b in range(len(nodes)):
        parts, ibv = block_values(b)
        payload = np.concatenate(parts).astype(endian + dtype).tobytes()
        if ibv is not None:
            payload += ibv.astype(endian + "i4").tobytes()
        chunks.append(framed(payload))
    with open(path, "wb") as f:
        f.write(b"".join(chunks))

"""Neutral Map File (NMF) connectivity and multiblock mesh assembly from Plot3D grids.

Follows the reference (`wmles/nmf.py:1-351`): `read_nmf`, `respan`, `extrude_2d`,
`build_mesh` (with sub-block splitting) and `describe`.

NMF faces (NASA GEOLAB convention): 1 kmin, 2 kmax, 3 imin, 4 imax, 5 jmin, 6 jmax, with
primary / secondary tangential axes (i, j), (i, j), (j, k), (j, k), (k, i), (k, i).
Ranges are 1-based inclusive node ranges; a ONE_TO_ONE partner's ranges may run
backwards (orientation), and Swap=TRUE pairs block 1's primary axis with block 2's
secondary axis.

`build_mesh` cuts every block along every patch boundary and carries the cuts across
one-to-one interfaces until no new cut appears, so that each face of each sub-block is
exactly one patch; the cut planes become interfaces, and every interface's orientation is
derived again from the node coordinates by `Mesh.connect`.
"""
from dataclasses import dataclass
from itertools import product

import numpy as np

from .bc import BC, NoSlipWall, SlipWall
from .mesh import Block, Interface, Mesh

# NMF face -> (normal axis, side, (primary axis, secondary axis))
FACES = {1: (2, 0, (0, 1)), 2: (2, 1, (0, 1)), 3: (0, 0, (1, 2)), 4: (0, 1, (1, 2)),
         5: (1, 0, (2, 0)), 6: (1, 1, (2, 0))}
FACE_NAME = {(0, 0): "imin", (0, 1): "imax", (1, 0): "jmin", (1, 1): "jmax", (2, 0): "kmin",
             (2, 1): "kmax"}
DEFAULT_BC_MAP = {"symmetry_x": SlipWall, "symmetry_y": SlipWall, "symmetry_z": SlipWall,
                  "wall": NoSlipWall, "viscous_solid": NoSlipWall, "inviscid_solid": SlipWall}
_UNSUPPORTED = ("collapsed", "pole_dir1", "pole_dir2", "patched", "unprocessed")
_TRUE = ("true", ".true.", "t", "1")


@dataclass
class Patch:
    type: str
    block: int                 # 0-based
    face: int                  # 1..6
    ranges: tuple              # ((S1, E1), (S2, E2)) as written (1-based)
    partner: object = None     # one_to_one: dict(block, face, ranges, swap)


@dataclass
class NMF:
    dims: list                 # node counts (ni, nj, nk) per block
    patches: list


def _norm_type(t):
    return t.lower().replace("-", "_")


def read_nmf(path):
    rows = []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].replace("\\", " ").strip()
            if line:
                rows.append(line.split())
    nb = int(rows[0][0])
    dims = [None] * nb
    for r in rows[1:1 + nb]:
        b = int(r[0])
        dims[b - 1] = tuple(int(t) for t in r[1:4])
    patches = []
    for r in rows[1 + nb:]:
        typ = _norm_type(r[0])
        if typ in _UNSUPPORTED:
            raise NotImplementedError(f"NMF entry type '{r[0]}' is not supported "
                                      "(only one-to-one interfaces and boundary conditions)")
        v = r[1:]
        if typ == "one_to_one":
            a = [int(t) for t in v[:12]]
            swap = len(v) > 12 and v[12].lower() in _TRUE
            patches.append(Patch(typ, a[0] - 1, a[1], ((a[2], a[3]), (a[4], a[5])),
                                 dict(block=a[6] - 1, face=a[7], ranges=((a[8], a[9]), (a[10], a[11])),
                                      swap=swap)))
        else:
            a = [int(t) for t in v[:6]]
            patches.append(Patch(typ, a[0] - 1, a[1], ((a[2], a[3]), (a[4], a[5]))))
    return NMF(dims, patches)


# ------------------------------------------------------------------------------ span
def respan(nodes, nmf, ncells, axis=0, length=None):
    """Blocks whose `axis` has exactly two node planes related by one common translation
    -> `ncells` uniform cells along it (span rescaled to `length` if given); the NMF node
    ranges along that axis are renumbered.  Inputs are not modified."""
    span = None
    for b, X in enumerate(nodes):
        if X.shape[1 + axis] != 2:
            raise ValueError(f"block {b + 1}: axis {axis} has {X.shape[1 + axis]} nodes, expected 2")
        d = np.take(X, 1, axis=1 + axis) - np.take(X, 0, axis=1 + axis)
        d0 = d.reshape(3, -1)[:, 0]
        if not np.allclose(d, d0.reshape((3,) + (1,) * (d.ndim - 1)), atol=1e-9 * max(np.abs(X).max(), 1.0)):
            raise ValueError(f"block {b + 1}: the two planes along axis {axis} are not a pure translation")
        if span is None:
            span = d0
        elif not np.allclose(d0, span, atol=1e-9 * max(np.linalg.norm(span), 1.0)):
            raise ValueError("blocks disagree on the span vector")
    step = span * (length / np.linalg.norm(span)) if length else span
    out = []
    for X in nodes:
        base = np.take(X, 0, axis=1 + axis)
        out.append(np.stack([base + (m / ncells) * step.reshape(3, 1, 1) for m in range(ncells + 1)],
                            axis=1 + axis))
    return out, _renumber(nmf, axis, 2, ncells + 1)


def _renumber(nmf, axis, old, new):
    dims = [tuple(new if a == axis else n for a, n in enumerate(d)) for d in nmf.dims]

    def fix(face, ranges):
        out = []
        for ax, (s, e) in zip(FACES[face][2], ranges):
            if ax == axis:
                if {s, e} != {1, old}:
                    raise ValueError(f"cannot respan: a patch covers only part of axis {axis}")
                s, e = (1, new) if s < e else (new, 1)
            out.append((s, e))
        return tuple(out)

    patches = []
    for p in nmf.patches:
        q = Patch(p.type, p.block, p.face, fix(p.face, p.ranges))
        if p.partner:
            q.partner = dict(p.partner, ranges=fix(p.partner["face"], p.partner["ranges"]))
        patches.append(q)
    return NMF(dims, patches)


def extrude_2d(nodes, ncells, length, direction=(0.0, 0.0, 1.0)):
    """(3, ni, nj, 1) blocks of a 2-D file -> (3, ni, nj, ncells + 1), extruded along
    `direction` over `length` (a new k axis; 3-D NMF files do not apply to the result)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d) * length
    return [np.stack([X[..., 0] + (m / ncells) * d.reshape(3, 1, 1) for m in range(ncells + 1)], axis=3)
            for X in nodes]


# ------------------------------------------------------------------------------ topology
def _tangential(face, ranges):
    """[(axis, lo, hi)] 0-based with lo <= hi for the two tangential axes of a patch."""
    return [(ax, min(s, e) - 1, max(s, e) - 1) for ax, (s, e) in zip(FACES[face][2], ranges)]


def _axis_pairs(p):
    """Per tangential axis of a one-to-one patch: (axis A, range A, axis B, range B),
    0-based, in the directions written in the file."""
    q = p.partner
    pa, sa = FACES[p.face][2]
    pb, sb = FACES[q["face"]][2]
    ra = [(s - 1, e - 1) for s, e in p.ranges]
    rb = [(s - 1, e - 1) for s, e in q["ranges"]]
    if q["swap"]:
        pairs = [(pa, ra[0], sb, rb[1]), (sa, ra[1], pb, rb[0])]
    else:
        pairs = [(pa, ra[0], pb, rb[0]), (sa, ra[1], sb, rb[1])]
    for _, (a0, a1), _, (b0, b1) in pairs:
        if abs(a1 - a0) != abs(b1 - b0):
            raise ValueError(f"inconsistent one_to_one ranges: block {p.block + 1} face {p.face} vs "
                             f"block {q['block'] + 1} face {q['face']}")
    return pairs


def _carry(m, ra, rb):
    """Node index m of range ra -> the corresponding index of range rb."""
    (a0, a1), (b0, b1) = ra, rb
    if a0 == a1:
        return b0
    return b0 + (m - a0) * (b1 - b0) // (a1 - a0)


def _bc_for(bc_map, typ):
    v = bc_map.get(typ, DEFAULT_BC_MAP.get(typ))
    if v is None:
        raise KeyError(f"no BC supplied for NMF type '{typ}' - pass bc_map={{'{typ}': <BC or factory>}}")
    return v if isinstance(v, BC) else v()


def _cut_planes(nmf):
    nb = len(nmf.dims)
    cuts = [[set(), set(), set()] for _ in range(nb)]
    for p in nmf.patches:
        sides = [(p.block, p.face, p.ranges)]
        if p.partner:
            sides.append((p.partner["block"], p.partner["face"], p.partner["ranges"]))
        for b, face, ranges in sides:
            for ax, lo, hi in _tangential(face, ranges):
                if lo > 0:
                    cuts[b][ax].add(lo)
                if hi < nmf.dims[b][ax] - 1:
                    cuts[b][ax].add(hi)
    links = [p for p in nmf.patches if p.partner]
    grew = True
    while grew:                                  # carry cuts across one-to-one faces
        grew = False
        for p in links:
            A, B = p.block, p.partner["block"]
            for ax_a, ra, ax_b, rb in _axis_pairs(p):
                for sb, sax, sr, db, dax, dr in ((A, ax_a, ra, B, ax_b, rb), (B, ax_b, rb, A, ax_a, ra)):
                    lo, hi = sorted(sr)
                    for m in sorted(c for c in cuts[sb][sax] if lo < c < hi):
                        m2 = _carry(m, sr, dr)
                        if m2 not in cuts[db][dax]:
                            cuts[db][dax].add(m2)
                            grew = True
    return cuts


def build_mesh(nodes, nmf, bc_map=None, periodic=(), ng=3, finalize=True):
    """A `Mesh` from Plot3D node arrays and their NMF, split into sub-blocks so that every
    sub-block face is exactly one patch.

    bc_map    {nmf type: BC instance or zero-argument factory}, over DEFAULT_BC_MAP
    periodic  (nmf type, axis) pairs: such patches on the faces normal to `axis` become a
              periodic connection between each sub-block's two faces (translation from
              the nodes)
    Each block gets `.origin = dict(block=<0-based NMF block>, ranges=((i0, i1), (j0, j1),
    (k0, k1)))`, the node range it covers in its NMF block."""
    bc_map = dict(bc_map or {})
    periodic = {(_norm_type(t), ax) for t, ax in periodic}
    nb = len(nodes)
    for b in range(nb):
        if tuple(nodes[b].shape[1:]) != tuple(nmf.dims[b]):
            raise ValueError(f"block {b + 1}: grid has {nodes[b].shape[1:]} nodes, NMF says {nmf.dims[b]}")
    cuts = _cut_planes(nmf)
    # sub-blocks, in block order then C order of their segment indices
    edges = [[[0] + sorted(cuts[b][ax]) + [nmf.dims[b][ax] - 1] for ax in range(3)] for b in range(nb)]
    subs, order = {}, []
    for b in range(nb):
        nseg = [len(edges[b][ax]) - 1 for ax in range(3)]
        for seg in product(*(range(n) for n in nseg)):
            rng = tuple((edges[b][ax][seg[ax]], edges[b][ax][seg[ax] + 1]) for ax in range(3))
            cells = tuple(r1 - r0 for r0, r1 in rng)
            if min(cells) < ng:
                raise ValueError(f"block {b + 1}{list(seg)} has only {cells} cells; every direction needs "
                                 f">= ng={ng} (a 1-cell-thick span must be widened first: see respan())")
            X = nodes[b][:, rng[0][0]:rng[0][1] + 1, rng[1][0]:rng[1][1] + 1, rng[2][0]:rng[2][1] + 1]
            name = f"b{b + 1}" if max(nseg) == 1 else f"b{b + 1}_" + "".join(str(s) for s in seg)
            blk = Block(name, X, ng=ng)
            blk.origin = dict(block=b, ranges=rng)
            subs[(b, seg)] = blk
            order.append((b, seg))
    mesh = Mesh([subs[k] for k in order])

    def on_face(key, ax, side):
        b = key[0]
        r = subs[key].origin["ranges"][ax]
        return r[side] == (0 if side == 0 else nmf.dims[b][ax] - 1)

    # cut planes -> interfaces between neighbouring sub-blocks of one block
    for (b, seg), blk in subs.items():
        for ax in range(3):
            nxt = seg[:ax] + (seg[ax] + 1,) + seg[ax + 1:]
            if (b, nxt) in subs:
                mesh.connect(blk, FACE_NAME[(ax, 1)], subs[(b, nxt)], FACE_NAME[(ax, 0)])
    # patches -> boundary conditions, one-to-one interfaces, periodic marks
    marks = set()
    for p in nmf.patches:
        ax, side, _ = FACES[p.face]
        tang = _tangential(p.face, p.ranges)
        for key, blk in subs.items():
            if key[0] != p.block or not on_face(key, ax, side):
                continue
            rng = blk.origin["ranges"]
            if not all(lo <= rng[a][0] and rng[a][1] <= hi for a, lo, hi in tang):
                continue
            fname = FACE_NAME[(ax, side)]
            if p.partner is None:
                if (p.type, ax) in periodic:
                    marks.add((key, ax, side))
                else:
                    blk.set_bc(fname, _bc_for(bc_map, p.type))
                continue
            if (ax, side) in blk.faces:          # joined already from the partner's entry
                continue
            q = p.partner
            bax, bside, _ = FACES[q["face"]]
            want = {}
            for ax_a, ra, ax_b, rb in _axis_pairs(p):
                lo, hi = rng[ax_a]
                want[ax_b] = tuple(sorted((_carry(lo, ra, rb), _carry(hi, ra, rb))))
            hits = [k for k, sb in subs.items() if k[0] == q["block"] and on_face(k, bax, bside)
                    and all(sb.origin["ranges"][a] == r for a, r in want.items())]
            if len(hits) != 1:
                raise ValueError(f"block {p.block + 1} face {p.face}: partner sub-face not found ({len(hits)})")
            mesh.connect(blk, fname, subs[hits[0]], FACE_NAME[(bax, bside)])
    # periodic pairs: both faces of one sub-block
    done = set()
    for key, ax, side in sorted(marks):
        if (key, ax) in done:
            continue
        if (key, ax, 1 - side) not in marks:
            raise ValueError(f"{subs[key].name}: periodic axis {ax} needs the periodic type on both faces")
        blk = subs[key]
        d = (np.take(blk.x_int, -1, axis=1 + ax) - np.take(blk.x_int, 0, axis=1 + ax)).reshape(3, -1)
        if not np.allclose(d, d[:, :1], atol=1e-9 * max(np.abs(blk.x_int).max(), 1.0)):
            raise ValueError(f"{blk.name}: axis {ax} faces are not related by a pure translation")
        mesh.connect(blk, FACE_NAME[(ax, 0)], blk, FACE_NAME[(ax, 1)], shift=tuple(-d[:, 0]))
        done.add((key, ax))
    if finalize:
        mesh.finalize()
    return mesh


def describe(mesh):
    """Text summary: blocks, sizes, origin, the BC or partner of every face, total cells."""
    lines, total = [], 0
    for i, b in enumerate(mesh.blocks):
        total += int(np.prod(b.n))
        o = getattr(b, "origin", None)
        src = f"  (block {o['block'] + 1}, nodes {[tuple(r) for r in o['ranges']]})" if o else ""
        lines.append(f"[{i}] {b.name}: cells {b.n}{src}")
        for (ax, s), f in sorted(b.faces.items()):
            if isinstance(f, Interface):
                other = mesh.blocks[f.nb]
                sh = "" if not np.any(f.shift) else f"  shift={[float(v) + 0.0 for v in np.round(f.shift, 6)]}"
                lines.append(f"      {FACE_NAME[(ax, s)]}: -> {other.name}.{FACE_NAME[tuple(f.src_face)]}{sh}")
            else:
                lines.append(f"      {FACE_NAME[(ax, s)]}: {type(f).__name__}")
    lines.append(f"total cells: {total}")
    return "\n".join(lines)

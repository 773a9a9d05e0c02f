"""Neutral Map File (NMF) connectivity and multiblock mesh assembly from Plot3D grids
(host side; the assembled `Mesh` runs through the unchanged HIP step).

    grid = read_plot3d("nak_a_ff_500c_2zn_L5.p3dfmt")
    nmf = read_nmf("nak_a_ff_500c_2zn_L5.nmf")
    nodes, nmf = respan(grid.nodes, nmf, ncells=8, length=0.1)
    mesh = build_mesh(nodes, nmf, bc_map={"viscous_solid": lambda: WallModel(...),
                                          "farfield_riem": Dirichlet(prim)},
                      periodic=(("symmetry_y", 0),))

Public names, arguments, the `NMF` / `Patch` fields, the sub-block names and the block
order are the reference's (`wmles/nmf.py:1-351`), since callers and tests see them.

The NMF format (NASA GEOLAB): a block count, one `block ni nj nk` line per block (node
counts), then one line per patch: `type block face S1 E1 S2 E2` for a boundary condition,
`one_to_one b1 f1 S1 E1 S2 E2 b2 f2 S1 E1 S2 E2 [swap]` for a block-to-block connection.
Faces 1..6 are kmin, kmax, imin, imax, jmin, jmax; ranges are 1-based inclusive node
ranges along the face's primary and secondary axes (i,j / j,k / k,i); on the second side
of a one-to-one entry a range may run backwards, and `swap` pairs the first side's
primary axis with the second side's secondary axis.  `#` starts a comment; a backslash is
layout only.

Assembly.  `Mesh.connect` joins whole faces, while an NMF patch may cover part of a face.
So every block is cut into sub-blocks at each patch boundary; a cut on one side of a
one-to-one patch is carried to the other side through the patch's index map, until no
new cut appears (a worklist; the closure does not depend on the visiting order).  Every
sub-block face then lies in exactly one patch: boundary patches give BCs, one-to-one
patches connect the sub-block to the partner sub-block whose face the index map hits, the
planes cut inside a block become ordinary interfaces, and `periodic` patch types become
translations between a sub-block's two faces.  `Mesh.connect` re-derives (and so checks)
every orientation from the node coordinates.
"""
from collections import deque
from dataclasses import dataclass
from itertools import product

import numpy as np

from .bc import BC, NoSlipWall, SlipWall
from .mesh import Interface, Mesh, Block

# NMF face number -> (normal axis, side, (primary axis, secondary axis)); fixed by the format
FACES = {1: (2, 0, (0, 1)), 2: (2, 1, (0, 1)),
         3: (0, 0, (1, 2)), 4: (0, 1, (1, 2)),
         5: (1, 0, (2, 0)), 6: (1, 1, (2, 0))}
FACE_NAME = {(a, s): "ijk"[a] + ("min", "max")[s] for a in range(3) for s in range(2)}

# boundary types with a default BC (classes: one instance per face); far-field, inflow
# and outflow types need an explicit entry in bc_map
DEFAULT_BC_MAP = {"symmetry_x": SlipWall, "symmetry_y": SlipWall, "symmetry_z": SlipWall,
                  "wall": NoSlipWall, "viscous_solid": NoSlipWall, "inviscid_solid": SlipWall}

_NOT_HANDLED = {"collapsed", "pole_dir1", "pole_dir2", "patched", "unprocessed"}


@dataclass
class Patch:
    type: str                # normalised: lower case, '-' -> '_'
    block: int               # 0-based
    face: int                # 1..6
    ranges: tuple            # ((S1, E1), (S2, E2)), 1-based inclusive, as written
    partner: object = None   # one-to-one: dict(block, face, ranges, swap)


@dataclass
class NMF:
    dims: list               # per block (ni, nj, nk) node counts
    patches: list


def _norm(t):
    return t.lower().replace("-", "_")


# ------------------------------------------------------------------------------ parsing
def read_nmf(path):
    """Parse an NMF file into `NMF(dims, patches)` (reference `wmles/nmf.py:78-107`)."""
    records = []
    with open(path) as f:
        for raw in f:
            toks = raw.partition("#")[0].replace("\\", " ").split()
            if toks:
                records.append(toks)
    if not records:
        raise ValueError(f"{path}: no NMF records")
    it = iter(records)
    nblocks = int(next(it)[0])
    dims = [None] * nblocks
    for _ in range(nblocks):
        rec = next(it, None)
        if rec is None or len(rec) < 4:
            raise ValueError(f"{path}: expected {nblocks} block dimension lines")
        b = int(rec[0]) - 1
        if not 0 <= b < nblocks or dims[b] is not None:
            raise ValueError(f"{path}: bad or repeated block number {rec[0]}")
        dims[b] = tuple(int(v) for v in rec[1:4])
    patches = []
    for rec in it:
        kind = _norm(rec[0])
        if kind in _NOT_HANDLED:
            raise NotImplementedError(f"{path}: NMF patch type '{rec[0]}' cannot be assembled; "
                                      "only boundary patches and one-to-one connections are")
        ints = [int(v) for v in rec[1:13 if kind == "one_to_one" else 7]]
        rng = ((ints[2], ints[3]), (ints[4], ints[5]))
        partner = None
        if kind == "one_to_one":
            if len(ints) < 12:
                raise ValueError(f"{path}: one_to_one entry with {len(ints)} of 12 integers")
            flag = rec[13].lower() if len(rec) > 13 else "false"
            partner = dict(block=ints[6] - 1, face=ints[7],
                           ranges=((ints[8], ints[9]), (ints[10], ints[11])),
                           swap=flag in ("true", ".true.", "t", "1"))
        elif len(ints) < 6:
            raise ValueError(f"{path}: '{rec[0]}' entry with {len(ints)} of 6 integers")
        patches.append(Patch(kind, ints[0] - 1, ints[1], rng, partner))
    return NMF(dims, patches)


# ------------------------------------------------------------------------------ span
def _span_vector(X, axis, label):
    """The translation between the two node planes of `axis` (which must have exactly 2)."""
    if X.shape[1 + axis] != 2:
        raise ValueError(f"{label}: {X.shape[1 + axis]} node planes along axis {axis}; respan needs 2")
    d = (np.take(X, 1, axis=1 + axis) - np.take(X, 0, axis=1 + axis)).reshape(3, -1)
    v = d[:, 0]
    scale = max(float(np.abs(X).max()), 1.0)
    if np.abs(d - v[:, None]).max() > 1e-9 * scale:
        raise ValueError(f"{label}: the planes along axis {axis} differ by more than a translation")
    return v


def respan(nodes, nmf, ncells, axis=0, length=None):
    """Replace a two-plane axis (the one-cell span of 2-D-derived grids such as the TMBWG
    airfoils) by `ncells` uniform cells, the total span optionally rescaled to `length`.
    Every block must share the span vector.  Returns (nodes, nmf); inputs untouched
    (reference `wmles/nmf.py:112-136`)."""
    span = None
    for b, X in enumerate(nodes):
        v = _span_vector(X, axis, f"block {b + 1}")
        if span is None:
            span = v
        elif np.abs(v - span).max() > 1e-9 * max(float(np.linalg.norm(span)), 1.0):
            raise ValueError(f"block {b + 1}: span vector {v} differs from block 1's {span}")
    if length:
        span = span * (length / np.linalg.norm(span))
    frac = np.arange(ncells + 1) / ncells
    out = []
    for X in nodes:
        base = np.take(X, 0, axis=1 + axis)
        planes = [base + f * span.reshape(3, 1, 1) for f in frac]
        out.append(np.stack(planes, axis=1 + axis))
    return out, _stretch_axis(nmf, axis, 2, ncells + 1)


def _stretch_axis(nmf, axis, old, new):
    """NMF with `axis` going from `old` to `new` nodes: every range along it must be the
    full axis (either direction)."""
    def conv(face, ranges):
        out = []
        for ax, (s, e) in zip(FACES[face][2], ranges):
            if ax == axis:
                if sorted((s, e)) != [1, old]:
                    raise ValueError(f"respan: a patch covers nodes {s}..{e} of axis {axis}, not all {old}")
                s, e = (1, new) if s <= e else (new, 1)
            out.append((s, e))
        return tuple(out)

    dims = [tuple(new if a == axis else n for a, n in enumerate(d)) for d in nmf.dims]
    patches = []
    for p in nmf.patches:
        partner = None
        if p.partner is not None:
            partner = dict(p.partner)
            partner["ranges"] = conv(partner["face"], partner["ranges"])
        patches.append(Patch(p.type, p.block, p.face, conv(p.face, p.ranges), partner))
    return NMF(dims, patches)


def extrude_2d(nodes, ncells, length, direction=(0.0, 0.0, 1.0)):
    """(3, ni, nj, 1) blocks of a 2-D grid -> (3, ni, nj, ncells + 1) blocks extruded by
    `length` along `direction` (the new k axis).  An NMF written for a 3-D grid does not
    describe the result (reference `wmles/nmf.py:139-145`)."""
    d = np.asarray(direction, float)
    d = d * (length / np.linalg.norm(d))
    frac = np.arange(ncells + 1) / ncells
    return [X[..., 0][..., None] + d.reshape(3, 1, 1, 1) * frac.reshape(1, 1, 1, -1) for X in nodes]


# ------------------------------------------------------------------------------ topology
class _Side:
    """One side of a patch: block, normal (axis, side), and per tangential axis its
    0-based node range (lo <= hi)."""

    def __init__(self, block, face, ranges):
        self.block = block
        self.axis, self.side, axes = FACES[face]
        self.span = {ax: (min(s, e) - 1, max(s, e) - 1) for ax, (s, e) in zip(axes, ranges)}

    def covers(self, sub_ranges):
        return all(lo <= sub_ranges[ax][0] and sub_ranges[ax][1] <= hi for ax, (lo, hi) in self.span.items())


class _IndexMap:
    """Node index map of one tangential axis across a one-to-one patch:
    idx_b = b0 + sign * (idx_a - a0) for idx_a in [min(a0, a1), max(a0, a1)]."""

    def __init__(self, ax_a, ra, ax_b, rb, where):
        a0, a1 = ra[0] - 1, ra[1] - 1
        b0, b1 = rb[0] - 1, rb[1] - 1
        if abs(a1 - a0) != abs(b1 - b0):
            raise ValueError(f"one_to_one {where}: ranges of {abs(a1 - a0)} and {abs(b1 - b0)} "
                             "node intervals cannot be matched")
        self.ax_a, self.ax_b = ax_a, ax_b
        self.a0, self.b0 = a0, b0
        self.sign = 1 if (a1 - a0) * (b1 - b0) >= 0 else -1
        self.lo, self.hi = min(a0, a1), max(a0, a1)

    def __call__(self, m):
        return self.b0 + self.sign * (m - self.a0)

    def inverse(self):
        inv = _IndexMap.__new__(_IndexMap)
        inv.ax_a, inv.ax_b, inv.a0, inv.b0, inv.sign = self.ax_b, self.ax_a, self.b0, self.a0, self.sign
        inv.lo, inv.hi = sorted((self(self.lo), self(self.hi)))
        return inv


def _maps_of(p):
    """The two tangential index maps (side A -> side B) of one-to-one patch p."""
    q = p.partner
    pa, sa = FACES[p.face][2]
    pb, sb = FACES[q["face"]][2]
    targets = (sb, pb) if q["swap"] else (pb, sb)
    tb_ranges = {pb: q["ranges"][0], sb: q["ranges"][1]}
    where = f"block {p.block + 1} face {p.face} / block {q['block'] + 1} face {q['face']}"
    return [_IndexMap(ax_a, ra, ax_b, tb_ranges[ax_b], where)
            for ax_a, ra, ax_b in zip((pa, sa), p.ranges, targets)]


def _bc_instance(bc_map, kind):
    entry = bc_map[kind] if kind in bc_map else DEFAULT_BC_MAP.get(kind)
    if entry is None:
        raise KeyError(f"NMF boundary type '{kind}' has no BC: add it to bc_map")
    return entry if isinstance(entry, BC) else entry()


def _cut_planes(nmf):
    """Per block and axis, the set of interior node planes where sub-blocks must meet."""
    nb = len(nmf.dims)
    cuts = [[set() for _ in range(3)] for _ in range(nb)]
    queue = deque()

    def add(b, ax, m):
        if 0 < m < nmf.dims[b][ax] - 1 and m not in cuts[b][ax]:
            cuts[b][ax].add(m)
            queue.append((b, ax, m))

    links = {}                          # (block, axis) -> [(map, partner block)]
    for p in nmf.patches:
        sides = [_Side(p.block, p.face, p.ranges)]
        if p.partner is not None:
            q = p.partner
            sides.append(_Side(q["block"], q["face"], q["ranges"]))
            for mp in _maps_of(p):
                links.setdefault((p.block, mp.ax_a), []).append((mp, q["block"]))
                links.setdefault((q["block"], mp.ax_b), []).append((mp.inverse(), p.block))
        for s in sides:
            for ax, (lo, hi) in s.span.items():
                add(s.block, ax, lo)
                add(s.block, ax, hi)
    while queue:
        b, ax, m = queue.popleft()
        for mp, other in links.get((b, ax), ()):
            if mp.lo < m < mp.hi:
                add(other, mp.ax_b, mp(m))
    return cuts


def build_mesh(nodes, nmf, bc_map=None, periodic=(), ng=3, finalize=True):
    """Assemble the `Mesh` of a Plot3D grid and its NMF (reference `wmles/nmf.py:206-332`).

    bc_map    {nmf type: BC instance or zero-argument factory}, over DEFAULT_BC_MAP
    periodic  (nmf type, axis) pairs: such patches on faces normal to `axis` become a
              periodic connection between a sub-block's two faces (translation from the
              nodes)
    Each block gets `.origin = dict(block=<0-based NMF block>, ranges=((i0, i1), (j0, j1),
    (k0, k1)))`, the node ranges it takes from the grid block."""
    bc_map = {_norm(k): v for k, v in (bc_map or {}).items()}
    periodic = {(_norm(t), int(ax)) for t, ax in periodic}
    nb = len(nodes)
    if len(nmf.dims) != nb:
        raise ValueError(f"the grid has {nb} blocks, the NMF {len(nmf.dims)}")
    for b in range(nb):
        if tuple(nodes[b].shape[1:]) != tuple(nmf.dims[b]):
            raise ValueError(f"block {b + 1}: grid nodes {tuple(nodes[b].shape[1:])} against NMF {nmf.dims[b]}")

    # sub-blocks, in grid-block order, then C order of their segment indices
    cuts = _cut_planes(nmf)
    edges = [[[0, *sorted(cuts[b][ax]), nmf.dims[b][ax] - 1] for ax in range(3)] for b in range(nb)]
    by_index, by_ranges, order = {}, {}, []
    for b in range(nb):
        nseg = [len(e) - 1 for e in edges[b]]
        for idx in product(*map(range, nseg)):
            rng = tuple((edges[b][ax][idx[ax]], edges[b][ax][idx[ax] + 1]) for ax in range(3))
            ncell = tuple(hi - lo for lo, hi in rng)
            if min(ncell) < ng:
                raise ValueError(f"block {b + 1}, sub-block {idx}: {ncell} cells, fewer than ng = {ng} "
                                 "in some direction (a one-cell span needs respan() first)")
            X = nodes[b][:, rng[0][0]:rng[0][1] + 1, rng[1][0]:rng[1][1] + 1, rng[2][0]:rng[2][1] + 1]
            name = f"b{b + 1}" if nseg == [1, 1, 1] else f"b{b + 1}_" + "".join(str(i) for i in idx)
            blk = Block(name, X, ng=ng)
            blk.origin = dict(block=b, ranges=rng)
            by_index[(b, idx)] = blk
            by_ranges[(b, rng)] = blk
            order.append(blk)
    mesh = Mesh(order)

    # planes cut inside a grid block
    for (b, idx), blk in by_index.items():
        for ax in range(3):
            up = tuple(i + (a == ax) for a, i in enumerate(idx))
            if (b, up) in by_index:
                mesh.connect(blk, FACE_NAME[(ax, 1)], by_index[(b, up)], FACE_NAME[(ax, 0)])

    # every patch side, with the index maps towards its partner side
    sides = []
    for p in nmf.patches:
        if p.partner is None:
            sides.append((_Side(p.block, p.face, p.ranges), p, None, None))
            continue
        q = p.partner
        maps = _maps_of(p)
        sides.append((_Side(p.block, p.face, p.ranges), p, q, maps))
        sides.append((_Side(q["block"], q["face"], q["ranges"]), p,
                      dict(block=p.block, face=p.face), [m.inverse() for m in maps]))

    periodic_faces = set()
    for blk in order:
        b, rng = blk.origin["block"], blk.origin["ranges"]
        for ax, side in product(range(3), range(2)):
            if rng[ax][side] != (0 if side == 0 else nmf.dims[b][ax] - 1):
                continue                                  # an inner plane: connected above
            owners = [s for s in sides if s[0].block == b and (s[0].axis, s[0].side) == (ax, side)
                      and s[0].covers(rng)]
            if not owners:
                continue                                  # left open: finalize() reports it
            if len(owners) > 1:
                raise ValueError(f"{blk.name}.{FACE_NAME[(ax, side)]} lies in {len(owners)} NMF patches")
            own, p, other, maps = owners[0]
            face = FACE_NAME[(ax, side)]
            if other is None:
                if (p.type, ax) in periodic:
                    periodic_faces.add((blk.name, ax, side))
                else:
                    blk.set_bc(face, _bc_instance(bc_map, p.type))
                continue
            if (ax, side) in blk.faces:
                continue                                  # joined from the partner's side
            pax, pside, _ = FACES[other["face"]]
            pb = other["block"]
            want = [None, None, None]
            for mp in maps:
                lo, hi = rng[mp.ax_a]
                want[mp.ax_b] = tuple(sorted((mp(lo), mp(hi))))
            pe = edges[pb][pax]
            want[pax] = (pe[0], pe[1]) if pside == 0 else (pe[-2], pe[-1])
            partner = by_ranges.get((pb, tuple(want)))
            if partner is None:
                raise ValueError(f"{blk.name}.{face}: no sub-block of grid block {pb + 1} has the "
                                 f"face its one-to-one patch maps to (node ranges {tuple(want)})")
            mesh.connect(blk, face, partner, FACE_NAME[(pax, pside)])

    # periodic translations between the two faces of one sub-block
    for blk in order:
        for ax in range(3):
            marks = [(blk.name, ax, s) in periodic_faces for s in range(2)]
            if not any(marks):
                continue
            if not all(marks):
                raise ValueError(f"{blk.name}: periodic type on only one {'ijk'[ax]} face")
            d = (np.take(blk.x_int, blk.n[ax], axis=1 + ax) - np.take(blk.x_int, 0, axis=1 + ax)).reshape(3, -1)
            if np.abs(d - d[:, :1]).max() > 1e-9 * max(float(np.abs(blk.x_int).max()), 1.0):
                raise ValueError(f"{blk.name}: its two {'ijk'[ax]} faces are not a translation apart")
            mesh.connect(blk, FACE_NAME[(ax, 0)], blk, FACE_NAME[(ax, 1)], shift=tuple(-d[:, 0]))

    if finalize:
        mesh.finalize()
    return mesh


def describe(mesh):
    """Text summary of an assembled mesh: each block's cells and origin, and what each face
    is joined to (an interface with its shift, or the BC class)."""
    out, total = [], 0
    for i, blk in enumerate(mesh.blocks):
        total += int(np.prod(blk.n))
        org = getattr(blk, "origin", None)
        where = "" if org is None else f" from grid block {org['block'] + 1} nodes {list(org['ranges'])}"
        out.append(f"{i:3d} {blk.name} {blk.n[0]}x{blk.n[1]}x{blk.n[2]} cells{where}")
        for key in sorted(blk.faces):
            f = blk.faces[key]
            if isinstance(f, Interface):
                tgt = f"{mesh.blocks[f.nb].name}.{FACE_NAME[tuple(f.src_face)]}"
                shift = "" if not np.any(f.shift) else " shift " + " ".join(f"{v:.6g}" for v in f.shift)
                out.append(f"      {FACE_NAME[key]} = {tgt}{shift}")
            else:
                out.append(f"      {FACE_NAME[key]} : {type(f).__name__}")
    out.append(f"cells: {total}")
    return "\n".join(out)

"""Structured multiblock mesh (oracle).  Follows reference `wmles/mesh.py`.

* ghosted blocks, cell arrays (..., Ni, Nj, Nk) with N = n + 2 ng      (mesh.py:88-121)
* geometry from ghost-extended nodes: face vectors S, face centres xf,
  divergence-theorem volume, centroid                                   (mesh.py:53-73)
* cell / node ghost exchange as an index remap                         (mesh.py:125-155)
* orientation detection by matching node planes                         (mesh.py:169-193)
* ghost-node construction by exchange / point reflection                (mesh.py:158-166,213-242)
"""
from itertools import product
from typing import NamedTuple

import numpy as np

FACE_NAMES = {"imin": (0, 0), "imax": (0, 1),
              "jmin": (1, 0), "jmax": (1, 1),
              "kmin": (2, 0), "kmax": (2, 1)}          # mesh.py:24-26


def sl(nd, d, s, other=slice(None)):
    """Index tuple: `s` on spatial axis d of an nd-array whose last 3 axes are
    spatial, `other` on the two remaining spatial axes (mesh.py:29-35)."""
    lead = nd - 3
    idx = [slice(None)] * lead + [other, other, other]
    idx[lead + d] = s
    return tuple(idx)


class Geom(NamedTuple):
    """mesh.py:38-43."""
    vol: np.ndarray
    xc: np.ndarray
    S: tuple
    xf: tuple


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1],
                     a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def _dot3(a, b):
    p = a * b
    return (p[0] + p[1]) + p[2]


def compute_geometry(X):
    """Metrics of a ghost-extended node array X (3, Ni+1, Nj+1, Nk+1) (mesh.py:53-73).

    For each direction the two in-face tangents are taken in cyclic order so S
    points along +d: S = 0.5 (c11 - c00) x (c01 - c10), xf = mean of 4 corners.
    """
    _, a1, a2, a3 = X.shape
    Ni, Nj, Nk = a1 - 1, a2 - 1, a3 - 1

    def quad(c00, c10, c01, c11):
        return 0.5 * _cross(c11 - c00, c01 - c10), 0.25 * (((c00 + c10) + c01) + c11)

    def corners(get):
        return get(0, 0), get(1, 0), get(0, 1), get(1, 1)

    Si, xi = quad(*corners(lambda a, b: X[:, :, a:Nj + a, b:Nk + b]))   # tangents (j, k)
    Sj, xj = quad(*corners(lambda a, b: X[:, b:Ni + b, :, a:Nk + a]))   # tangents (k, i)
    Sk, xk = quad(*corners(lambda a, b: X[:, a:Ni + a, b:Nj + b, :]))   # tangents (i, j)
    vol = (_dot3(xi[:, 1:], Si[:, 1:]) - _dot3(xi[:, :-1], Si[:, :-1])
           + _dot3(xj[:, :, 1:], Sj[:, :, 1:]) - _dot3(xj[:, :, :-1], Sj[:, :, :-1])
           + _dot3(xk[..., 1:], Sk[..., 1:]) - _dot3(xk[..., :-1], Sk[..., :-1])) / 3.0
    acc = 0
    for a, b, c in product((0, 1), repeat=3):
        acc = acc + X[:, a:Ni + a, b:Nj + b, c:Nk + c]
    xc = acc / 8.0
    return vol, xc, (Si, Sj, Sk), (xi, xj, xk)


class Interface:
    """Connection seen from the destination block (mesh.py:77-85)."""

    def __init__(self, nb, src_face, mapping, flip, shift):
        self.nb = nb
        self.src_face = src_face
        self.mapping = tuple(mapping)
        self.flip = tuple(flip)
        self.shift = np.asarray(shift, float)


class Block:
    """Ghosted structured block (mesh.py:88-121)."""

    def __init__(self, name, nodes, ng=3):
        nodes = np.asarray(nodes, float)
        if nodes.ndim != 4 or nodes.shape[0] != 3:
            raise ValueError("nodes must be (3, ni+1, nj+1, nk+1)")
        self.name, self.ng, self.x_int = name, ng, nodes
        self.n = tuple(s - 1 for s in nodes.shape[1:])
        if min(self.n) < ng:
            raise ValueError("each block needs at least ng cells per direction")
        self.N = tuple(k + 2 * ng for k in self.n)
        self.faces = {}
        self.X = None
        self.geom = None

    def set_bc(self, face, bc):
        self.faces[FACE_NAMES[face]] = bc

    def interior(self, a):
        lead, ng = a.ndim - 3, self.ng
        return a[(slice(None),) * lead + tuple(slice(ng, ng + k) for k in self.n)]

    def take(self, a, axis, idx):
        lead, ng = a.ndim - 3, self.ng
        t = [slice(None)] * lead + [slice(ng, ng + k) for k in self.n]
        t[lead + axis] = idx
        return a[tuple(t)]

    def slab(self, a, d, lo, hi):
        lead, ng = a.ndim - 3, self.ng
        t = [slice(None)] * lead + [slice(ng, ng + k) for k in self.n]
        t[lead + d] = slice(ng + lo, ng + self.n[d] + hi)
        return a[tuple(t)]


def exchange(dst, src, blkA, faceA, blkB, itf, kind):
    """Return `dst` with its ghost layers on faceA filled from `src` (mesh.py:125-155).

    cell: ghost layer m <- m-th cell of B off the interface;
    node: the interface node plane is shared and `itf.shift` is added.
    """
    aA, sA = faceA
    aB, sB = itf.src_face
    ng, nA, nB = blkA.ng, blkA.n[aA], blkB.n[aB]
    lead = dst.ndim - 3
    if kind == "cell":
        lo_src = ng if sB == 0 else nB
        lo_dst = ng + nA if sA == 1 else 0
    else:
        lo_src = ng + 1 if sB == 0 else nB
        lo_dst = ng + nA + 1 if sA == 1 else 0
    blk = src[sl(src.ndim, aB, slice(lo_src, lo_src + ng))]
    if sB == 1:
        blk = np.flip(blk, axis=lead + aB)                 # to off-interface order m
    blk = np.transpose(blk, list(range(lead)) + [lead + itf.mapping[t] for t in range(3)])
    for t in range(3):
        if itf.flip[t]:
            blk = np.flip(blk, axis=lead + t)
    if kind == "node":
        blk = blk + itf.shift.reshape(3, 1, 1, 1)
    if sA == 0:
        blk = np.flip(blk, axis=lead + aA)                 # m order -> position order
    out = dst.copy()
    out[sl(dst.ndim, aA, slice(lo_dst, lo_dst + ng))] = blk
    return out


def _reflect_nodes(X, blk, axis, side):
    """Point reflection of nodes through the boundary plane, 2 X_b - X_mirror
    (mesh.py:158-166).  In place."""
    ng, n = blk.ng, blk.n[axis]
    if side == 0:
        mirror = np.flip(X[sl(4, axis, slice(ng + 1, ng + 1 + ng))], axis=1 + axis)
        X[sl(4, axis, slice(0, ng))] = 2 * X[sl(4, axis, slice(ng, ng + 1))] - mirror
    else:
        b = ng + n
        mirror = np.flip(X[sl(4, axis, slice(n, n + ng))], axis=1 + axis)
        X[sl(4, axis, slice(b + 1, b + 1 + ng))] = 2 * X[sl(4, axis, slice(b, b + 1))] - mirror


def _match_orientation(A, fa, B, fb, shift):
    """Axis mapping / tangential flips with B's interface node plane + shift == A's
    (mesh.py:169-193): 2 tangent permutations x 4 flip combinations, allclose with
    atol = 1e-9 * max|plane_A|."""
    aA, sA = fa
    aB, sB = fb

    def plane(blk, a, s):
        return np.take(blk.x_int, 0 if s == 0 else blk.n[a], axis=1 + a)

    pa = plane(A, aA, sA)
    pb = plane(B, aB, sB) + np.asarray(shift, float).reshape(3, 1, 1)
    tA = [m for m in range(3) if m != aA]
    tB = [m for m in range(3) if m != aB]
    scale = max(np.abs(pa).max(), 1e-30)
    for perm in ([tB[0], tB[1]], [tB[1], tB[0]]):
        cand0 = np.transpose(pb, [0] + [1 + tB.index(p) for p in perm])
        for f1, f2 in product((0, 1), repeat=2):
            c = cand0
            if f1:
                c = np.flip(c, 1)
            if f2:
                c = np.flip(c, 2)
            if c.shape == pa.shape and np.allclose(c, pa, atol=1e-9 * scale, rtol=0):
                mapping, flip = [0, 0, 0], [False, False, False]
                mapping[aA], mapping[tA[0]], mapping[tA[1]] = aB, perm[0], perm[1]
                flip[tA[0]], flip[tA[1]] = bool(f1), bool(f2)
                return tuple(mapping), tuple(flip)
    raise ValueError(f"interface {A.name}{fa} <-> {B.name}{fb}: node planes do not match")


class Mesh:
    """mesh.py:196-242."""

    def __init__(self, blocks):
        self.blocks = list(blocks)
        if len({b.ng for b in self.blocks}) != 1:
            raise ValueError("all blocks must share ng")
        self.ng = self.blocks[0].ng
        self.geoms = None

    def index(self, blk):
        return next(i for i, b in enumerate(self.blocks) if b is blk)

    def connect(self, A, faceA, B, faceB, shift=(0.0, 0.0, 0.0)):
        fa, fb = FACE_NAMES[faceA], FACE_NAMES[faceB]
        shift = np.asarray(shift, float)
        for X, fx, Y, fy, sh in ((A, fa, B, fb, shift), (B, fb, A, fa, -shift)):
            mapping, flip = _match_orientation(X, fx, Y, fy, sh)
            X.faces[fx] = Interface(self.index(Y), fy, mapping, flip, sh)

    def finalize(self, node_passes=3):
        ng = self.ng
        for b in self.blocks:
            for f in FACE_NAMES.values():
                if f not in b.faces:
                    raise KeyError(f"block {b.name}: face {f} has no BC / connection")
            X = np.zeros((3,) + tuple(k + 1 + 2 * ng for k in b.n))
            X[(slice(None),) + tuple(slice(ng, ng + k + 1) for k in b.n)] = b.x_int
            b.X = X
        for _ in range(node_passes):
            for a in range(3):
                for b in self.blocks:
                    for s in (0, 1):
                        f = b.faces[(a, s)]
                        if isinstance(f, Interface):
                            b.X = exchange(b.X, self.blocks[f.nb].X, b, (a, s),
                                           self.blocks[f.nb], f, "node")
                        else:
                            _reflect_nodes(b.X, b, a, s)
        geoms = []
        for b in self.blocks:
            vol, xc, S, xf = compute_geometry(b.X)
            if vol.min() <= 0:
                raise ValueError(f"block {b.name}: non-positive cell volume")
            b.geom = Geom(vol, xc, S, xf)
            geoms.append(b.geom)
        self.geoms = tuple(geoms)
        return self

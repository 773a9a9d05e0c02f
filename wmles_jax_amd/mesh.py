"""Structured multiblock mesh on the HIP device (reference `wmles/mesh.py`).

Same interface as the reference: `Block(name, nodes, ng=3)`, `Mesh(blocks)`,
`Mesh.connect(A, faceA, B, faceB, shift)`, `Mesh.finalize(node_passes=3)`, `Geom`.
Orientation detection compares the two interface node planes on the host (setup
logic, reference `mesh.py:169-193`); ghost nodes and the metric terms are then built
on the device by `wmx_fill_nodes` / `wmx_compute_geometry`, and `Geom` holds torch
tensors in the reference layout.
"""
from itertools import product
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .backend import device

FACE_NAMES = {"imin": (0, 0), "imax": (0, 1),
              "jmin": (1, 0), "jmax": (1, 1),
              "kmin": (2, 0), "kmax": (2, 1)}


def sl(nd, d, s, other=slice(None)):
    """Index tuple for an nd-array whose last 3 axes are spatial (mesh.py:29-35)."""
    lead = nd - 3
    t = [slice(None)] * lead + [other] * 3
    t[lead + d] = s
    return tuple(t)


class Geom(NamedTuple):
    """Block metrics (mesh.py:38-43): vol (Ni,Nj,Nk), xc (3,..), S[d], xf[d] (3,..N_d+1..)."""
    vol: torch.Tensor
    xc: torch.Tensor
    S: tuple
    xf: tuple


class Interface:
    """Block-to-block connection seen from the destination block (mesh.py:77-85)."""

    def __init__(self, nb, src_face, mapping, flip, shift):
        self.nb = nb
        self.src_face = src_face
        self.mapping = tuple(mapping)
        self.flip = tuple(flip)
        self.shift = np.asarray(shift, float)

    def _fill_desc(self, fd, ctx=None):
        fd.kind = _lib.FACE_INTERFACE
        fd.nb = self.nb
        fd.src_axis, fd.src_side = self.src_face
        for t in range(3):
            fd.mapping[t] = self.mapping[t]
            fd.flip[t] = int(self.flip[t])
            fd.shift[t] = float(self.shift[t])


class Block:
    """Ghosted structured block (mesh.py:88-121)."""

    def __init__(self, name, nodes, ng=3):
        nodes = np.ascontiguousarray(nodes, dtype=float)    # views (transposes, flips) copied
        if nodes.ndim != 4 or nodes.shape[0] != 3:
            raise ValueError("nodes must be (3, ni+1, nj+1, nk+1)")
        if ng != 3:
            raise ValueError("the HIP step is built for ng = 3 ghost layers")
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


def _match_orientation(A, fa, B, fb, shift):
    """Axis mapping and tangential flips for which B's interface node plane plus
    `shift` equals A's: 2 tangent permutations x 4 flip combinations, tolerance
    1e-9 max|plane_A| (mesh.py:169-193)."""
    aA, sA = fa
    aB, sB = fb

    def plane(blk, a, s):
        return np.take(blk.x_int, 0 if s == 0 else blk.n[a], axis=1 + a)

    pa = plane(A, aA, sA)
    pb = plane(B, aB, sB) + np.asarray(shift, float).reshape(3, 1, 1)
    tA = [m for m in range(3) if m != aA]
    tB = [m for m in range(3) if m != aB]
    tol = 1e-9 * max(np.abs(pa).max(), 1e-30)
    for perm in ([tB[0], tB[1]], [tB[1], tB[0]]):
        base = np.transpose(pb, [0] + [1 + tB.index(p) for p in perm])
        for f1, f2 in product((0, 1), repeat=2):
            c = base
            if f1:
                c = np.flip(c, 1)
            if f2:
                c = np.flip(c, 2)
            if c.shape == pa.shape and np.allclose(c, pa, atol=tol, rtol=0):
                mapping, flip = [0, 0, 0], [False, False, False]
                mapping[aA], mapping[tA[0]], mapping[tA[1]] = aB, perm[0], perm[1]
                flip[tA[0]], flip[tA[1]] = bool(f1), bool(f2)
                return tuple(mapping), tuple(flip)
    raise ValueError(f"interface {A.name}{fa} <-> {B.name}{fb}: node planes do not match "
                     f"(check the shift vector and the block node coordinates)")


def block_desc(blk, index_of, ctx=None, geom=None, case_id=0):
    """C descriptor of one block: dims, faces, geometry pointers."""
    d = _lib.BlockDesc()
    for a in range(3):
        d.n[a] = blk.n[a]
    d.case_id = case_id
    for (a, s), f in blk.faces.items():
        fd = d.faces[2 * a + s]
        fd.inflow_id = -1
        fd.T_wall = -1.0
        fd.n_match = 1
        f._fill_desc(fd, ctx)
    if geom is not None:
        d.vol = geom.vol.data_ptr()
        d.xc = geom.xc.data_ptr()
        for a in range(3):
            d.S[a] = geom.S[a].data_ptr()
            d.xf[a] = geom.xf[a].data_ptr()
    return d


class _NodeCtx:
    """Face descriptors for the node sweep need only interface vs boundary."""
    nodes_only = True


class Mesh:
    """mesh.py:196-242."""

    def __init__(self, blocks):
        self.blocks = list(blocks)
        if len({b.ng for b in self.blocks}) != 1:
            raise ValueError("all blocks must share ng")
        if len(self.blocks) > _lib.WMX_MAX_BLOCKS:
            raise ValueError(f"at most {_lib.WMX_MAX_BLOCKS} blocks per mesh")
        self.ng = self.blocks[0].ng
        self.geoms = None

    def index(self, blk):
        return next(i for i, b in enumerate(self.blocks) if b is blk)

    def connect(self, A, faceA, B, faceB, shift=(0.0, 0.0, 0.0)):
        """Join A.faceA to B.faceB; `shift` = x_A(plane) - x_B(plane)."""
        fa, fb = FACE_NAMES[faceA], FACE_NAMES[faceB]
        shift = np.asarray(shift, float)
        for X, fx, Y, fy, sh in ((A, fa, B, fb, shift), (B, fb, A, fa, -shift)):
            mapping, flip = _match_orientation(X, fx, Y, fy, sh)
            X.faces[fx] = Interface(self.index(Y), fy, mapping, flip, sh)

    def finalize(self, node_passes=3):
        ng = self.ng
        dev = device()
        for b in self.blocks:
            for f in FACE_NAMES.values():
                if f not in b.faces:
                    raise KeyError(f"block {b.name}: face {f} has no BC / connection")
        Xs = []
        for b in self.blocks:
            X = torch.zeros((3,) + tuple(k + 1 + 2 * ng for k in b.n), dtype=torch.float64, device=dev)
            X[(slice(None),) + tuple(slice(ng, ng + k + 1) for k in b.n)] = torch.as_tensor(
                b.x_int, dtype=torch.float64, device=dev)
            b.X = X
            Xs.append(X)
        descs = (_lib.BlockDesc * len(self.blocks))(
            *[block_desc(b, self.index, _NodeCtx()) for b in self.blocks])
        _lib.check(_lib.lib.wmx_fill_nodes(len(self.blocks), descs, _lib.ptr_array(Xs), ng,
                                            node_passes, _lib.stream_ptr()))
        geoms = []
        for b in self.blocks:
            N = b.N
            opt = dict(dtype=torch.float64, device=dev)
            vol = torch.empty(N, **opt)
            xc = torch.empty((3,) + N, **opt)
            S = tuple(torch.empty((3,) + tuple(N[t] + (t == d) for t in range(3)), **opt)
                      for d in range(3))
            xf = tuple(torch.empty_like(s) for s in S)
            dims = (_lib.C.c_int32 * 3)(*N)
            _lib.check(_lib.lib.wmx_compute_geometry(
                b.X.data_ptr(), dims, vol.data_ptr(), xc.data_ptr(), S[0].data_ptr(),
                S[1].data_ptr(), S[2].data_ptr(), xf[0].data_ptr(), xf[1].data_ptr(),
                xf[2].data_ptr(), _lib.stream_ptr()))
            if float(vol.min()) <= 0:
                raise ValueError(f"block {b.name}: non-positive cell volume (left-handed or folded grid)")
            b.geom = Geom(vol, xc, S, xf)
            geoms.append(b.geom)
        self.geoms = tuple(geoms)
        return self

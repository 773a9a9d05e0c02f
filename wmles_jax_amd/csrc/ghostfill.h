// Per-cell ghost values of the sequential sweep (reference solver.py:59-73, mesh.py:125-155,
// bc.py:14-143), shared by the fill kernel (ghost.hip) and the primitive kernels that evaluate
// the sweep's last phase themselves (reduce.hip).  Both units are built with
// -ffp-contract=off, so both give the same bits.
#pragma once
#include "common.h"

namespace wmx {

// Decode thread -> (m, c_dst[3]) for the ghost slab of (axis, side); dims = array extents.
__device__ __forceinline__ void ghost_coords(int idx, int axis, int ng, const int* dims, int* m,
                                             int* c) {
  int t1, t2;
  tangents(axis, &t1, &t2);
  if (axis == 2) {                 // k faces: m fastest (contiguous along k)
    *m = idx % ng;
    int r = idx / ng;
    c[1] = r % dims[1];
    c[0] = r / dims[1];
  } else {
    c[t2] = idx % dims[t2];
    int r = idx / dims[t2];
    c[t1] = r % dims[t1];
    *m = r / dims[t1];
  }
}

// Cell / node exchange, reference wmles/mesh.py:125-155.  node != 0: node arrays
// (extents N+1, interface plane shared, shift added).
__device__ __forceinline__ void interface_src(const wmx_face_desc& f, int axis, int side, int m,
                                              const int* c_dst, const int* dims_dst, int ng,
                                              int nB, int node, int* c_src) {
  int aB = f.src_axis, sB = f.src_side;
  for (int t = 0; t < 3; ++t) {
    if (t == axis) continue;
    int v = f.flip[t] ? dims_dst[t] - 1 - c_dst[t] : c_dst[t];
    c_src[f.mapping[t]] = v;
  }
  if (node)
    c_src[aB] = sB == 1 ? nB + ng - 1 - m : ng + 1 + m;
  else
    c_src[aB] = sB == 1 ? nB + ng - 1 - m : ng + m;
}

// Value of ghost cell c (c[a] already the ghost coordinate, m its layer off the face) of
// block blk for face (a, s), into out[5].  false for inflow faces (k_inflow fills them).
__device__ __forceinline__ bool ghost_value(const BlockDev* __restrict__ blks, int blk, int a,
                                            int s, int m, const int* c, const Ptrs& q, const GasC& gas,
                                            double* out) {
  const BlockDev& B = blks[blk];
  const int ng = B.ng, n = B.n[a];
  const wmx_face_desc& f = B.faces[2 * a + s];
  if (f.kind == WMX_BC_INFLOW) return false;
  if (f.kind == WMX_FACE_INTERFACE) {
    const BlockDev& S = blks[f.nb];
    int cs[3];
    interface_src(f, a, s, m, c, B.N, ng, S.n[f.src_axis], 0, cs);
    const double* qs = q.p[f.nb];
    const long long si = cidx(S.N, cs[0], cs[1], cs[2]);
    for (int v = 0; v < 5; ++v) out[v] = qs[v * S.cells + si];
    return true;
  }
  const double* qd = q.p[blk];
  const long long cells = B.cells;
  int ci[3] = {c[0], c[1], c[2]};
  ci[a] = s == 0 ? ng + m : ng + n - 1 - m;
  if (f.kind == WMX_BC_EXTRAPOLATE) ci[a] = s == 0 ? ng : ng + n - 1;
  const long long ii = cidx(B.N, ci[0], ci[1], ci[2]);
  double src[5];
  for (int v = 0; v < 5; ++v) src[v] = qd[v * cells + ii];
  if (f.kind == WMX_BC_EXTRAPOLATE) {
    for (int v = 0; v < 5; ++v) out[v] = src[v];
  } else if (f.kind == WMX_BC_DIRICHLET) {
    prim_to_cons(gas, f.prim[0], f.prim[1], f.prim[2], f.prim[3], f.prim[4], out);
  } else if (f.kind == WMX_BC_NOSLIP) {
    bc_noslip(gas, src, f.T_wall, out);
  } else {
    // unit normal of the boundary face at this tangential position (bc.py:17-19)
    int cf[3] = {c[0], c[1], c[2]};
    cf[a] = s == 0 ? ng : ng + n;
    const long long fi = fidx(B.N, a, cf[0], cf[1], cf[2]);
    const long long fc = fcount(B.N, a);
    const double* S = B.S[a];
    double Sx = S[fi], Sy = S[fc + fi], Sz = S[2 * fc + fi];
    double Sm = sqrt((Sx * Sx + Sy * Sy) + Sz * Sz);
    double nx = Sx / Sm, ny = Sy / Sm, nz = Sz / Sm;
    if (f.kind == WMX_BC_FARFIELD)
      bc_farfield(gas, src, f.prim, nx, ny, nz, s, out);
    else                                            // SLIP, WALLMODEL
      bc_slip(gas, src, nx, ny, nz, out);
  }
  return true;
}

// Ghost cell idx of fill op `op` (one thread per ghost cell of the op's slab).
__device__ __forceinline__ void fill_cell(const BlockDev* __restrict__ blks, const GhostOp& op, const Ptrs& q,
                                          const GasC& gas, int idx) {
  if (idx >= op.count) return;
  const BlockDev& B = blks[op.blk];
  const int ng = B.ng;
  int m, c[3];
  ghost_coords(idx, op.axis, ng, B.N, &m, c);
  const int a = op.axis, s = op.side, n = B.n[a];
  c[a] = s == 1 ? ng + n + m : ng - 1 - m;
  double out[5];
  if (!ghost_value(blks, op.blk, a, s, m, c, q, gas, out)) return;   // inflow: k_inflow
  double* qd = q.p[op.blk];
  const long long cells = B.cells;
  const long long di = cidx(B.N, c[0], c[1], c[2]);
  for (int v = 0; v < 5; ++v) qd[v * cells + di] = out[v];
}

}  // namespace wmx

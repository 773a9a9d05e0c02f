// Primitive-record workgroup body (prim32_block) with the CFL partial, the fused k-face
// ghost fill and the sponge-mean head, shared by the primitive kernels (reduce.hip) and the
// persistent stage kernel (stage_off.hip); the unit that includes it is built without FP
// contraction, so the CFL partials equal k_dt_partial's bit for bit.
#pragma once
#include "ghostfill.h"
#include "wallface.h"

namespace wmx {

// block of workgroup wg: the largest b < nblk with prefix[b] <= wg (prefix non-decreasing,
// prefix[0] = 0), by bisection (an ensemble plan has up to 64 blocks)
__device__ __forceinline__ int find_block_r(const int* __restrict__ prefix, int nblk, int wg) {
  int lo = 0, hi = nblk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= wg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// NaN-propagating min (numpy.min semantics)
__device__ __forceinline__ double nmin(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}

template <bool MIN>
__device__ double block_reduce(double v) {
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o, 64);
    v = MIN ? nmin(v, u) : v + u;
  }
  if (lane == 0) red[wave] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  double r = red[0];
  for (int w = 1; w < nw; ++w) r = MIN ? nmin(r, red[w]) : r + red[w];
  __syncthreads();
  return r;
}

// interior cell l of block B -> ghosted linear index (32-bit: plans keep every plane
// below 2^29 cells)
__device__ __forceinline__ unsigned interior_cell(const BlockDev& B, unsigned l, int* I, int* J,
                                                  int* K) {
  const unsigned n1 = (unsigned)B.n[1], n2 = (unsigned)B.n[2], ng = (unsigned)B.ng;
  const unsigned r = l / n2;
  *K = (int)(ng + (l - r * n2));
  *J = (int)(ng + r % n1);
  *I = (int)(ng + r / n1);
  return ((unsigned)*I * (unsigned)B.N[1] + (unsigned)*J) * (unsigned)B.N[2] + (unsigned)*K;
}

// CFL metric of cell (I, J, K): the cell means of its face vectors
// Sb[d] = 0.5 (S_lo + S_hi) in m[3 d + x] and their norms |Sb[d]| in m[9 + d]
// (solver.py:145-146).  The per-cell geometry part of the CFL dt; k_cell_sb stores the
// nine means at workspace binding (the norms are recomputed from them: 72 B per cell read
// instead of 96), and cell_dt evaluates everything itself when the plan has no such store
// (the same arithmetic in the same unit, so the same bits).
constexpr int NSB = 12;
static_assert(NSB_PLANES == 9, "workspace planes of the CFL metric: the nine cell-mean components");
__device__ __forceinline__ void metric_norms(double* m) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    m[9 + d] = sqrt((m[3 * d] * m[3 * d] + m[3 * d + 1] * m[3 * d + 1]) + m[3 * d + 2] * m[3 * d + 2]);
}
__device__ __forceinline__ void cell_metric(const BlockDev& B, int I, int J, int K, double* m) {
  const int cc[3] = {I, J, K};
  double lo[3][3], hi[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* S = B.S[d];
    const unsigned fc = (unsigned)fcount(B.N, d);
    int up[3] = {cc[0], cc[1], cc[2]};
    up[d] += 1;
    const unsigned f0 = (unsigned)fidx(B.N, d, cc[0], cc[1], cc[2]);
    const unsigned f1 = (unsigned)fidx(B.N, d, up[0], up[1], up[2]);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      lo[d][x] = ldu(S + x * fc, f0);
      hi[d][x] = ldu(S + x * fc, f1);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int x = 0; x < 3; ++x) m[3 * d + x] = 0.5 * (lo[d][x] + hi[d][x]);
  metric_norms(m);
}

// cell_dt's operands from the stored metric (planes B.Sb) and the volume of cell c.  A plane
// that holds +0.0 in every interior cell (B.sb_zero, found at workspace binding) is not read:
// the operand is that same +0.0, so the arithmetic and its bits are unchanged.  On an
// axis-aligned block this skips six of the nine planes (48 of the 80 bytes per cell).
__device__ __forceinline__ void cell_dt_load(const BlockDev& B, unsigned c, double* m, double* vol) {
  const size_t cells = (size_t)B.cells;
  const unsigned zero = (unsigned)B.sb_zero;
  *vol = ldu(B.vol, c);
#pragma unroll
  for (int i = 0; i < NSB_PLANES; ++i) {
    if ((zero >> i) & 1u) m[i] = 0.0;
    else m[i] = ldu(B.Sb + i * cells, c);
  }
  metric_norms(m);
}

// CFL time step of a cell from its primitives w, metric m and volume; solver.py:139-150, in
// the reference's association order (bit-exact)
__device__ __forceinline__ double cell_dt_eval(const NumDev& nd, const Prim& w, const double* m, double vol) {
  const GasC& gas = nd.gas;
  const double cs = sqrt(gas.gamma * gas.R * w.T);
  const double cnu = fmax(4.0 / 3.0, gas.gamma / gas.Pr);
  const double nu = (cnu * mu_suth(gas, w.T)) / w.r;
  double lam = 0.0, lamv = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double Sm = m[9 + d];
    lam = (lam + fabs(dot3(w.u, w.v, w.w, m[3 * d], m[3 * d + 1], m[3 * d + 2]))) + cs * Sm;
    lamv = lamv + (nu * (Sm * Sm)) / vol;
  }
  return nd.cfl / ((lam + 4.0 * lamv) / vol);
}

// CFL time step of interior cell (I, J, K) (ghosted index c) from its primitives w.  Every
// load first (one memory round trip), then the arithmetic.  The dt kernel and the fused
// stage-1 primitive kernels below share it in this unit, so the two give the same bits.
__device__ __forceinline__ double cell_dt(const BlockDev& B, const NumDev& nd, unsigned c, int I,
                                          int J, int K, const Prim& w) {
  double m[NSB], vol;
  if (B.Sb) {                                    // stored at workspace binding (k_cell_sb)
    cell_dt_load(B, c, m, &vol);
  } else {
    vol = ldu(B.vol, c);
    cell_metric(B, I, J, K, m);
  }
  return cell_dt_eval(nd, w, m, vol);
}

// Spanwise mean of sponge row `row` for rows of 9..129 cells (numpy's pairwise order for
// the elements 1..nk-1 by eight lanes u of the row, combined by shuffles in that order);
// every lane of the 8-lane group calls it.  smooth_ramp.py:214-216.
__device__ __forceinline__ void sponge_row_fast(const BlockDev& B, long long row, int u, bool live,
                                                long long c0, const double* __restrict__ qb) {
  const long long rows = (long long)B.n[0] * B.n[1];
  const int n = B.n[2] - 1;
  const double* vol = B.vol + c0;
  const long long cells = B.cells;
  double a[6];
  const int nfull = n - n % 8;
  if (live) {
    const double vv = vol[1 + u];
    a[0] = vv;
#pragma unroll
    for (int v = 0; v < 5; ++v) a[1 + v] = qb[v * cells + c0 + 1 + u] * vv;
    for (int i = 8; i < nfull; i += 8) {
      const double w = vol[1 + i + u];
      a[0] = a[0] + w;
#pragma unroll
      for (int v = 0; v < 5; ++v) a[1 + v] = a[1 + v] + qb[v * cells + c0 + 1 + i + u] * w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < 6; ++v) a[v] = 0.0;
  }
#pragma unroll
  for (int m = 1; m <= 4; m <<= 1)
#pragma unroll
    for (int v = 0; v < 6; ++v) a[v] = a[v] + __shfl_xor(a[v], m, 8);
  if (!live || u != 0) return;
  for (int i = nfull; i < n; ++i) {
    const double w = vol[1 + i];
    a[0] = a[0] + w;
    for (int v = 0; v < 5; ++v) a[1 + v] = a[1 + v] + qb[v * cells + c0 + 1 + i] * w;
  }
  const double den = vol[0] + a[0];
  for (int v = 0; v < 5; ++v) B.sponge_mean[v * rows + row] = (qb[v * cells + c0] * vol[0] + a[1 + v]) / den;
}

// Head workgroups of a primitive launch (a multiple of 8, so the primitive workgroups keep
// their XCD spans): workgroup h < sh.nwg takes 32 sponge rows of the block whose range
// sh.pre[b] <= h < sh.pre[b + 1] holds it (rows of 9..129 cells only).
__device__ __forceinline__ void sponge_head(const BlockDev* __restrict__ blks, int nblk, const Ptrs& q,
                                            const SpongeHead& sh, int h) {
  if (h >= sh.nwg) return;
  const int b = find_block_r(sh.pre, nblk, h);
  const BlockDev& B = blks[b];
  const long long rows = (long long)B.n[0] * B.n[1];
  const long long row = (long long)(h - sh.pre[b]) * 32 + (threadIdx.x >> 3);
  const int u = threadIdx.x & 7;
  const bool live = row < rows && B.sponge_row[row];
  const int ng = B.ng;
  const long long c0 = live ? cidx(B.N, ng + (int)(row / B.n[1]), ng + (int)(row % B.n[1]), ng) : 0;
  sponge_row_fast(B, row, u, live, c0, q.p[b]);
}

// ------------------------------------------------------------------------------ prim
// Primitive records of the cells that are read (rec_read: ghosts up to 2 layers deep).  DT:
// the stage-1 launch of a step
// also takes the CFL partial of its workgroup's interior cells into part[workgroup],
// reusing the state it loads (the per-case min then follows as for k_dt_partial).
__device__ __forceinline__ void cell_ijk(const BlockDev& B, unsigned c, int* I, int* J, int* K) {
  const unsigned N1 = (unsigned)B.N[1], N2 = (unsigned)B.N[2];
  const unsigned r = c / N2;
  *K = (int)(c - r * N2);
  *J = (int)(r % N1);
  *I = (int)(r / N1);
}
__device__ __forceinline__ bool interior_of(const BlockDev& B, unsigned c, int* I, int* J, int* K) {
  cell_ijk(B, c, I, J, K);
  const int ng = B.ng;
  return *I >= ng && *I < ng + B.n[0] && *J >= ng && *J < ng + B.n[1] && *K >= ng &&
         *K < ng + B.n[2];
}
// Cells whose primitive record some kernel reads: the block's cells and the two ghost
// layers next to them (the face stencils and the gradient records reach 2 cells past the
// block's cells; the gradient cells, 1).  The outer ghost layers (ng > 2) keep their
// state, but no record is formed for them.
__device__ __forceinline__ bool rec_read(const BlockDev& B, int I, int J, int K) {
  const int lo = B.ng - 2;
  return I >= lo && I < B.N[0] - lo && J >= lo && J < B.N[1] - lo && K >= lo && K < B.N[2] - lo;
}

// State of cell c: from q, or (zfuse) for a k ghost of a flagged side the value the sweep's
// last phase would store there (ghost_value, the fill kernel's own code)
__device__ __forceinline__ void prim_state(const BlockDev* __restrict__ blks, int b, long long c,
                                           const Ptrs& q, const GasC& gas, int zfuse, double* qv) {
  const BlockDev& B = blks[b];
  if (zfuse && B.zfuse) {
    // 32-bit decode (every block holds fewer than 2^32 cells); I, J only for the ghosts
    const unsigned N2 = (unsigned)B.N[2];
    const unsigned r = (unsigned)c / N2;
    const int K = (int)((unsigned)c - r * N2), ng = B.ng, n2 = B.n[2];
    const int side = K < ng ? 0 : (K >= ng + n2 ? 1 : -1);
    if (side >= 0 && ((B.zfuse >> side) & 1)) {
      const unsigned N1 = (unsigned)B.N[1];
      const int cc[3] = {(int)(r / N1), (int)(r % N1), K};
      const int m = side == 0 ? ng - 1 - K : K - ng - n2;
      if (ghost_value(blks, b, 2, side, m, cc, q, gas, qv)) return;
    }
  }
  const double* qb = q.p[b];
  const long long cells = B.cells;
#pragma unroll
  for (int v = 0; v < 5; ++v) qv[v] = qb[v * cells + c];
}

// FP32-bulk record: prims evaluated in FP64, stored in FP32 with p, T and e_int
// shifted by the block gauge so their small variations keep FP32 precision.
// Workgroup wgx of nwg of one primitive launch; called by every thread of a 256-thread
// workgroup (the k_prim32 kernel, and the persistent stage kernel per workgroup index).
template <bool DT>
__device__ __forceinline__ void prim32_block(const BlockDev* __restrict__ blks,
                                             const int* __restrict__ wg_prefix, int nblk,
                                             const Ptrs& q, const NumDev& nd, double* __restrict__ part,
                                             int zfuse, const SpongeHead& sh, int wgx, int nwg) {
  if (wgx < sh.head) {
    if (wgx < sh.nwg) {
      sponge_head(blks, nblk, q, sh, wgx);
    } else if (sh.wall_wgx > 0) {                 // Newton wall faces (reads interior q only)
      const int h = wgx - sh.nwg, op_i = h / sh.wall_wgx;
      if (op_i < sh.nwall) {
        const WallOp op = sh.wall[op_i];
        const int idx0 = (h - op_i * sh.wall_wgx) * (int)blockDim.x + (int)threadIdx.x;
        if (idx0 < op.count) wall_face<false>(blks, op, idx0, q, nd, nullptr, nullptr);
      }
    }
    return;
  }
  // XCD-contiguous spans: the i faces one plane apart that the CFL partial reads twice
  // stay in one L2
  const int wg = xcd_span(wgx - sh.head, nwg - sh.head);
  const int b = find_block_r(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  const long long c = (long long)(wg - wg_prefix[b]) * blockDim.x + threadIdx.x;
  const long long cells = B.cells;
  double dtc = __builtin_huge_val();
  int I = 0, J = 0, K = 0;
  if (c < cells) cell_ijk(B, (unsigned)c, &I, &J, &K);
  if (c < cells && rec_read(B, I, J, K)) {
    const GasC& gas = nd.gas;
    double qv[5];
    prim_state(blks, b, c, q, gas, zfuse, qv);
    rec32_store(B.Pf, (size_t)cells, (size_t)c, nd, qv, B.p_ref, B.T_gauge, B.e_ref);
    if (DT && interior_of(B, (unsigned)c, &I, &J, &K)) {
      const Prim w = cons_to_prim(gas, qv[0], qv[1], qv[2], qv[3], qv[4]);
      dtc = cell_dt(B, nd, (unsigned)c, I, J, K, w);
    }
  }
  if (DT) {
    const double r = block_reduce<true>(dtc);
    if (threadIdx.x == 0) part[wg] = r;
  }
}

}  // namespace wmx

// Whole-domain reductions of the step: CFL time step (solver.py:135-154), the
// constant-mass-flux constraint (solver.py:16-36) and the sponge's spanwise mean
// (smooth_ramp.py:214-216).  Two-level: per-workgroup partials (wavefront shuffles
// + LDS), then one workgroup per ensemble case.  Also the primitive kernels
// (gas.py:24-33), whose stage-1 launch of a step can take the CFL partials, so the
// separate dt kernel and the fused one evaluate the same code.  Built with
// -ffp-contract=off.
#include "ghostfill.h"

namespace wmx {

__device__ __forceinline__ int find_block_r(const int* __restrict__ prefix, int nblk, int wg) {
  int b = 0;
  while (b + 1 < nblk && prefix[b + 1] <= wg) ++b;
  return b;
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

// CFL time step of interior cell (I, J, K) (ghosted index c) from its primitives w;
// solver.py:139-150.  Every load first (one memory round trip), then the arithmetic.  The
// dt kernel and the fused stage-1 primitive kernels below share it in this unit, so the
// two give the same bits.
__device__ __forceinline__ double cell_dt(const BlockDev& B, const NumDev& nd, unsigned c, int I,
                                          int J, int K, const Prim& w) {
  double Sv[3][3][2];
  const double vol = ldu(B.vol, c);
  const int cc[3] = {I, J, K};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* S = B.S[d];
    const unsigned fc = (unsigned)fcount(B.N, d);
    int hi[3] = {cc[0], cc[1], cc[2]};
    hi[d] += 1;
    const unsigned f0 = (unsigned)fidx(B.N, d, cc[0], cc[1], cc[2]);
    const unsigned f1 = (unsigned)fidx(B.N, d, hi[0], hi[1], hi[2]);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      Sv[d][x][0] = ldu(S + x * fc, f0);
      Sv[d][x][1] = ldu(S + x * fc, f1);
    }
  }
  const GasC& gas = nd.gas;
  const double cs = sqrt(gas.gamma * gas.R * w.T);
  const double cnu = fmax(4.0 / 3.0, gas.gamma / gas.Pr);
  const double nu = (cnu * mu_suth(gas, w.T)) / w.r;
  double lam = 0.0, lamv = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double s0 = 0.5 * (Sv[d][0][0] + Sv[d][0][1]);
    const double s1 = 0.5 * (Sv[d][1][0] + Sv[d][1][1]);
    const double s2 = 0.5 * (Sv[d][2][0] + Sv[d][2][1]);
    const double Sm = sqrt((s0 * s0 + s1 * s1) + s2 * s2);
    lam = (lam + fabs(dot3(w.u, w.v, w.w, s0, s1, s2))) + cs * Sm;
    lamv = lamv + (nu * (Sm * Sm)) / vol;
  }
  return nd.cfl / ((lam + 4.0 * lamv) / vol);
}

// Per-cell CFL time step, min over interiors; solver.py:139-150.
__global__ __launch_bounds__(256) void k_dt_partial(const BlockDev* __restrict__ blks,
                                                    const int* __restrict__ wg_prefix, int nblk,
                                                    Ptrs q, NumDev nd, double* __restrict__ part) {
  const int wg = xcd_span(blockIdx.x, gridDim.x);
  const int b = find_block_r(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  // INT_CELLS interior cells per thread, blockDim.x apart
  const unsigned base = (unsigned)(wg - wg_prefix[b]) * (INT_CELLS * blockDim.x) + threadIdx.x;
  const unsigned nint = (unsigned)B.n[0] * B.n[1] * B.n[2];
  double dtc = __builtin_huge_val();
  for (int u = 0; u < INT_CELLS; ++u) {
    const unsigned l = base + (unsigned)u * blockDim.x;
    if (l < nint) {
      int I, J, K;
      const unsigned c = interior_cell(B, l, &I, &J, &K);
      const unsigned cells = (unsigned)B.cells;
      const double* qb = q.p[b];
      double qv[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) qv[v] = ldu(qb + v * cells, c);
      const Prim w = cons_to_prim(nd.gas, qv[0], qv[1], qv[2], qv[3], qv[4]);
      dtc = nmin(dtc, cell_dt(B, nd, c, I, J, K, w));
    }
  }
  const double r = block_reduce<true>(dtc);
  if (threadIdx.x == 0) part[wg] = r;
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
__device__ __forceinline__ void sponge_head(const BlockDev* __restrict__ blks, int nblk, Ptrs q,
                                            const SpongeHead& sh) {
  const int h = (int)blockIdx.x;
  if (h >= sh.nwg) return;
  int b = 0;
  while (b + 1 < nblk && sh.pre[b + 1] <= h) ++b;
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
// Primitive records of every cell (ghosts included).  DT: the stage-1 launch of a step
// also takes the CFL partial of its workgroup's interior cells into part[workgroup],
// reusing the state it loads (the per-case min then follows as for k_dt_partial).
__device__ __forceinline__ bool interior_of(const BlockDev& B, unsigned c, int* I, int* J, int* K) {
  const unsigned N1 = (unsigned)B.N[1], N2 = (unsigned)B.N[2];
  const unsigned r = c / N2;
  *K = (int)(c - r * N2);
  *J = (int)(r % N1);
  *I = (int)(r / N1);
  const int ng = B.ng;
  return *I >= ng && *I < ng + B.n[0] && *J >= ng && *J < ng + B.n[1] && *K >= ng &&
         *K < ng + B.n[2];
}

// State of cell c: from q, or (zfuse) for a k ghost of a flagged side the value the sweep's
// last phase would store there (ghost_value, the fill kernel's own code)
__device__ __forceinline__ void prim_state(const BlockDev* __restrict__ blks, int b, long long c,
                                           Ptrs q, const GasC& gas, int zfuse, double* qv) {
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

// FP64 record (r, u, v, w, p, T, mu, e_int)
template <bool DT>
__global__ __launch_bounds__(256) void k_prim(const BlockDev* __restrict__ blks,
                                              const int* __restrict__ wg_prefix, int nblk,
                                              Ptrs q, NumDev nd, double* __restrict__ part,
                                              int zfuse, SpongeHead sh) {
  if ((int)blockIdx.x < sh.head) {
    sponge_head(blks, nblk, q, sh);
    return;
  }
  // XCD-contiguous spans: the i faces one plane apart that the CFL partial reads twice
  // stay in one L2
  const int wg = xcd_span((int)blockIdx.x - sh.head, (int)gridDim.x - sh.head);
  const int b = find_block_r(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  const long long c = (long long)(wg - wg_prefix[b]) * blockDim.x + threadIdx.x;
  const long long cells = B.cells;
  double dtc = __builtin_huge_val();
  if (c < cells) {
    const GasC& gas = nd.gas;
    double qv[5];
    prim_state(blks, b, c, q, gas, zfuse, qv);
    const Prim w = cons_to_prim(gas, qv[0], qv[1], qv[2], qv[3], qv[4]);
    double* P = B.P;
    P[c] = w.r;
    P[cells + c] = w.u;
    P[2 * cells + c] = w.v;
    P[3 * cells + c] = w.w;
    P[4 * cells + c] = w.p;
    P[5 * cells + c] = w.T;
    P[6 * cells + c] = mu_suth(gas, w.T);
    P[7 * cells + c] = w.p / ((gas.gamma - 1.0) * w.r);
    int I, J, K;
    if (DT && interior_of(B, (unsigned)c, &I, &J, &K)) dtc = cell_dt(B, nd, (unsigned)c, I, J, K, w);
  }
  if (DT) {
    const double r = block_reduce<true>(dtc);
    if (threadIdx.x == 0) part[wg] = r;
  }
}

// FP32-bulk record: prims evaluated in FP64, stored in FP32 with p, T and e_int
// shifted by the block gauge so their small variations keep FP32 precision.
template <bool DT>
__global__ __launch_bounds__(256) void k_prim32(const BlockDev* __restrict__ blks,
                                                const int* __restrict__ wg_prefix, int nblk,
                                                Ptrs q, NumDev nd, double* __restrict__ part,
                                                int zfuse, SpongeHead sh) {
  if ((int)blockIdx.x < sh.head) {
    sponge_head(blks, nblk, q, sh);
    return;
  }
  // XCD-contiguous spans: the i faces one plane apart that the CFL partial reads twice
  // stay in one L2
  const int wg = xcd_span((int)blockIdx.x - sh.head, (int)gridDim.x - sh.head);
  const int b = find_block_r(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  const long long c = (long long)(wg - wg_prefix[b]) * blockDim.x + threadIdx.x;
  const long long cells = B.cells;
  double dtc = __builtin_huge_val();
  if (c < cells) {
    const GasC& gas = nd.gas;
    double qv[5];
    prim_state(blks, b, c, q, gas, zfuse, qv);
    const Prim w = cons_to_prim(gas, qv[0], qv[1], qv[2], qv[3], qv[4]);
    st4(B.Pf + 4 * (size_t)c, make_float4((float)w.u, (float)w.v, (float)w.w, (float)(w.T - B.T_gauge)));
    st4(B.Pf + 4 * ((size_t)cells + c),
        make_float4((float)w.r, (float)(w.p - B.p_ref), (float)mu_suth(gas, w.T),
                    (float)(w.p / ((gas.gamma - 1.0) * w.r) - B.e_ref)));
    int I, J, K;
    if (DT && interior_of(B, (unsigned)c, &I, &J, &K)) dtc = cell_dt(B, nd, (unsigned)c, I, J, K, w);
  }
  if (DT) {
    const double r = block_reduce<true>(dtc);
    if (threadIdx.x == 0) part[wg] = r;
  }
}

// sum over interiors of rho u_dir vol (and vol); solver.py:23-27
__global__ __launch_bounds__(256) void k_mf_partial(const BlockDev* __restrict__ blks,
                                                    const int* __restrict__ wg_prefix, int nblk,
                                                    Ptrs q, const int* __restrict__ case_dir,
                                                    double* __restrict__ part) {
  const int b = find_block_r(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  // INT_CELLS interior cells per thread, blockDim.x apart, added in that order
  const unsigned base = (unsigned)(blockIdx.x - wg_prefix[b]) * (INT_CELLS * blockDim.x) + threadIdx.x;
  const unsigned nint = (unsigned)B.n[0] * B.n[1] * B.n[2];
  const int dir = case_dir[B.case_id];
  double num = 0.0, den = 0.0;
  if (dir >= 0) {
    double vv[INT_CELLS], qm[INT_CELLS];
#pragma unroll
    for (int u = 0; u < INT_CELLS; ++u) {
      const unsigned l = base + (unsigned)u * blockDim.x;
      vv[u] = 0.0;
      qm[u] = 0.0;
      if (l < nint) {
        int I, J, K;
        const unsigned c = interior_cell(B, l, &I, &J, &K);
        vv[u] = ldu(B.vol, c);
        qm[u] = ldu(q.p[b] + (unsigned)(1 + dir) * (unsigned)B.cells, c);
      }
    }
#pragma unroll
    for (int u = 0; u < INT_CELLS; ++u) {
      num = num + qm[u] * vv[u];
      den = den + vv[u];
    }
  }
  const double rn = block_reduce<false>(num);
  const double rd = block_reduce<false>(den);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rn;
    part[2 * blockIdx.x + 1] = rd;
  }
}

// One workgroup (1024 threads) per case: combine the partials of the case's
// workgroups, four independent loads in flight per thread.
// MODE 0: dt = min;  MODE 1: delta = target - num / den
template <int MODE>
__global__ __launch_bounds__(1024) void k_case_final(const int* __restrict__ case_wg,
                                                     const int* __restrict__ case_wg_off,
                                                     const double* __restrict__ part,
                                                     const double* __restrict__ target,
                                                     double* __restrict__ outv) {
  const int cs = blockIdx.x;
  const int lo = case_wg_off[cs], hi = case_wg_off[cs + 1];
  double a = MODE == 0 ? __builtin_huge_val() : 0.0, b = 0.0;
  const int stride = blockDim.x;
  // case_wg == nullptr: the list is the identity (blocks ordered by case), so the partial
  // index is the list position and all eight loads of an iteration are independent
  for (int t0 = lo + threadIdx.x; t0 < hi; t0 += 8 * stride) {
    double va[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * stride;
      const bool ok = t < hi;
      const int wg = ok ? (case_wg ? case_wg[t] : t) : 0;
      va[u] = ok ? (MODE == 0 ? part[wg] : part[2 * wg]) : (MODE == 0 ? __builtin_huge_val() : 0.0);
      vb[u] = (ok && MODE == 1) ? part[2 * wg + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (MODE == 0) a = nmin(a, va[u]);
      else { a += va[u]; b += vb[u]; }
    }
  }
  if (MODE == 0) {
    a = block_reduce<true>(a);
    if (threadIdx.x == 0) outv[cs] = a;
  } else {
    a = block_reduce<false>(a);
    b = block_reduce<false>(b);
    if (threadIdx.x == 0) outv[cs] = target[cs] - a / b;
  }
}

template __global__ void k_case_final<0>(const int*, const int*, const double*, const double*, double*);
template __global__ void k_case_final<1>(const int*, const int*, const double*, const double*, double*);

// q_dir += delta; rho E += (2 (rho u_dir) delta + delta^2) / (2 rho); all cells (solver.py:28-36)
__global__ __launch_bounds__(256) void k_mf_apply(const BlockDev* __restrict__ blks,
                                                  const int* __restrict__ wg_prefix, int nblk,
                                                  Ptrs q, const int* __restrict__ case_dir,
                                                  const double* __restrict__ delta) {
  const int b = find_block_r(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  const long long c = (long long)(blockIdx.x - wg_prefix[b]) * blockDim.x + threadIdx.x;
  if (c >= B.cells) return;
  const int dir = case_dir[B.case_id];
  if (dir < 0) return;
  const double dl = delta[B.case_id];
  double* qb = q.p[b];
  const long long cells = B.cells;
  const double qm = qb[(1 + dir) * cells + c];
  const double q0 = qb[c];
  const double dE = ((2.0 * qm) * dl + dl * dl) / (2.0 * q0);
  // only rho u_dir and rho E change (the reference adds 0 to the other components,
  // which can only turn a -0.0 into +0.0)
  qb[(1 + dir) * cells + c] = qm + (0.0 + dl);
  qb[4 * cells + c] = qb[4 * cells + c] + dE;
}

// Plane statistics of one block (wmles/channel5200.py:139-157): per interior j plane the
// (i, k) means of rho, u, v, w, uu, vv, ww, uv, uw, vw, T, p.  Partial sums per
// (plane, chunk) workgroup, then one thread per plane sums its chunks in order and adds
// the means to the accumulator acc[NSTAT][n1].
__global__ __launch_bounds__(256) void k_plane_partial(const BlockDev* __restrict__ blks, int blk,
                                                       Ptrs q, GasC gas, int chunk,
                                                       double* __restrict__ part) {
  const BlockDev& B = blks[blk];
  const int j = blockIdx.y, ch = blockIdx.x, nch = gridDim.x;
  const int n0 = B.n[0], n2 = B.n[2], ng = B.ng;
  const long long cells = B.cells;
  const long long tot = (long long)n0 * n2;
  const long long l0 = (long long)ch * chunk, l1 = l0 + chunk < tot ? l0 + chunk : tot;
  const double* qb = q.p[blk];
  double a[NSTAT];
#pragma unroll
  for (int v = 0; v < NSTAT; ++v) a[v] = 0.0;
  for (long long l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
    const int i = (int)(l / n2), k = (int)(l % n2);
    const long long c = cidx(B.N, ng + i, ng + j, ng + k);
    const Prim w = cons_to_prim(gas, qb[c], qb[cells + c], qb[2 * cells + c], qb[3 * cells + c],
                                qb[4 * cells + c]);
    a[0] += w.r;
    a[1] += w.u;
    a[2] += w.v;
    a[3] += w.w;
    a[4] += w.u * w.u;
    a[5] += w.v * w.v;
    a[6] += w.w * w.w;
    a[7] += w.u * w.v;
    a[8] += w.u * w.w;
    a[9] += w.v * w.w;
    a[10] += w.T;
    a[11] += w.p;
  }
  for (int v = 0; v < NSTAT; ++v) {
    const double r = block_reduce<false>(a[v]);
    if (threadIdx.x == 0) part[((long long)j * nch + ch) * NSTAT + v] = r;
  }
}

__global__ void k_plane_final(int n1, int nch, double inv_count, const double* __restrict__ part,
                              double* __restrict__ acc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n1) return;
  for (int v = 0; v < NSTAT; ++v) {
    double s = 0.0;
    for (int ch = 0; ch < nch; ++ch) s += part[((long long)j * nch + ch) * NSTAT + v];
    acc[(long long)v * n1 + j] += s * inv_count;
  }
}

// Rows that need the sponge mean: some cell of the (i, j) row has a weight > 0 under
// some sponge (the weight depends on the static cell centre only; computed once).
__global__ void k_sponge_rows(const BlockDev* __restrict__ blks, int blk, NumDev nd) {
  const BlockDev& B = blks[blk];
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long rows = (long long)B.n[0] * B.n[1];
  if (row >= rows) return;
  const int i = (int)(row / B.n[1]), j = (int)(row % B.n[1]);
  const int ng = B.ng;
  const long long c0 = cidx(B.N, ng + i, ng + j, ng);
  int need = 0;
  for (int k = 0; k < B.n[2] && !need; ++k)
    for (int sp = 0; sp < nd.n_sponge; ++sp)
      if (sponge_sigma(nd.sponge[sp], B.xc[c0 + k]) != 0.0) need = 1;
  B.sponge_row[row] = need;
}

// Volume-weighted spanwise (k) mean of the interior state per (i, j) row
// (smooth_ramp.py:214-216) for the rows that need it, in numpy's add.reduce order: the
// first element plus the pairwise sum of the rest, whose runs of <= 128 use 8 interleaved
// accumulators.  Eight lanes take one row: lane u holds accumulator u, the three
// combining levels are lane exchanges in the same association, lane 0 adds the tail.
__global__ __launch_bounds__(256) void k_sponge_mean(const BlockDev* __restrict__ blks, int blk, Ptrs q) {
  const BlockDev& B = blks[blk];
  const long long rows = (long long)B.n[0] * B.n[1];
  const long long row = (long long)blockIdx.x * 32 + (threadIdx.x >> 3);
  const int u = threadIdx.x & 7;
  const bool live = row < rows && B.sponge_row[row];
  const int ng = B.ng, nk = B.n[2];
  const long long c0 = live ? cidx(B.N, ng + (int)(row / B.n[1]), ng + (int)(row % B.n[1]), ng) : 0;
  const double* vol = B.vol + c0;
  const double* qb = q.p[blk];
  const long long cells = B.cells;
  const int n = nk - 1;                                  // the pairwise part: elements 1..nk-1
  if (n <= 128 && n >= 8) {
    sponge_row_fast(B, row, u, live, c0, qb);
    return;
  }
  if (!live || u != 0) return;                           // short or long rows: one lane per row
  const double den = np_sum([&](int k) { return vol[k]; }, nk);
  for (int v = 0; v < 5; ++v) {
    const double* qv = qb + v * cells + c0;
    const double num = np_sum([&](int k) { return qv[k] * vol[k]; }, nk);
    B.sponge_mean[v * rows + row] = num / den;
  }
}

// ----------------------------------------------------------------------- launchers
hipError_t launch_dt_partial(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             const NumDev& nd, double* d_part, hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dt_partial, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q, nd, d_part);
  return hipGetLastError();
}

hipError_t launch_prim(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                       const NumDev& nd, double* d_part, int zfuse, SpongeHead sh, hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  const bool f32 = nd.prec == WMX_PREC_FP32_BULK;
  sh.head = sh.pre ? ((sh.nwg + 7) / 8) * 8 : 0;
  if (!sh.pre) sh.nwg = 0;
  const dim3 grid(nwg + sh.head);
  if (d_part) {
    if (f32) hipLaunchKernelGGL(k_prim32<true>, grid, dim3(256), 0, s, d_blks, d_pre, nblk, q, nd, d_part, zfuse, sh);
    else hipLaunchKernelGGL(k_prim<true>, grid, dim3(256), 0, s, d_blks, d_pre, nblk, q, nd, d_part, zfuse, sh);
  } else {
    if (f32) hipLaunchKernelGGL(k_prim32<false>, grid, dim3(256), 0, s, d_blks, d_pre, nblk, q, nd, d_part, zfuse, sh);
    else hipLaunchKernelGGL(k_prim<false>, grid, dim3(256), 0, s, d_blks, d_pre, nblk, q, nd, d_part, zfuse, sh);
  }
  return hipGetLastError();
}

hipError_t launch_plane_stats(const BlockDev* d_blks, const BlockDev& hb, int blk, Ptrs q,
                              const GasC& gas, double* d_part, double* d_acc, hipStream_t s) {
  const long long tot = (long long)hb.n[0] * hb.n[2];
  const int chunk = (int)((tot + PLANE_CHUNKS - 1) / PLANE_CHUNKS);
  hipLaunchKernelGGL(k_plane_partial, dim3(PLANE_CHUNKS, hb.n[1]), dim3(256), 0, s, d_blks, blk, q, gas,
                     chunk, d_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_plane_final, dim3((hb.n[1] + 63) / 64), dim3(64), 0, s, hb.n[1], PLANE_CHUNKS,
                     1.0 / (double)tot, d_part, d_acc);
  return hipGetLastError();
}

hipError_t launch_mf_partial(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             const int* d_case_dir, double* d_part, hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mf_partial, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q, d_case_dir,
                     d_part);
  return hipGetLastError();
}

hipError_t launch_case_final(int mode, int ncases, const int* d_case_wg, const int* d_case_off,
                             const double* d_part, const double* d_target, double* d_out,
                             hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL(k_case_final<0>, dim3(ncases), dim3(1024), 0, s, d_case_wg, d_case_off, d_part,
                       d_target, d_out);
  else
    hipLaunchKernelGGL(k_case_final<1>, dim3(ncases), dim3(1024), 0, s, d_case_wg, d_case_off, d_part,
                       d_target, d_out);
  return hipGetLastError();
}

hipError_t launch_mf_apply(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                           const int* d_case_dir, const double* d_delta, hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mf_apply, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q, d_case_dir,
                     d_delta);
  return hipGetLastError();
}

hipError_t launch_sponge_mean(const BlockDev* d_blks, int blk, long long rows, Ptrs q,
                              hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sponge_mean, dim3((unsigned)((rows + 31) / 32)), dim3(256), 0, s, d_blks, blk, q);
  return hipGetLastError();
}

hipError_t launch_sponge_rows(const BlockDev* d_blks, int blk, long long rows, const NumDev& nd,
                              hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sponge_rows, dim3((unsigned)((rows + 127) / 128)), dim3(128), 0, s, d_blks, blk, nd);
  return hipGetLastError();
}

}  // namespace wmx

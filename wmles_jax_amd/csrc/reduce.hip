// Whole-domain reductions of the step: CFL time step (solver.py:135-154), the
// constant-mass-flux constraint (solver.py:16-36) and the sponge's spanwise mean
// (smooth_ramp.py:214-216).  Two-level: per-workgroup partials (wavefront shuffles
// + LDS), then one workgroup per ensemble case.  Also the primitive kernels
// (gas.py:24-33), whose stage-1 launch of a step can take the CFL partials, so the
// separate dt kernel and the fused one evaluate the same code.  Built with
// -ffp-contract=off.
#include "prim32.h"

namespace wmx {

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

// FP64 record (r, u, v, w, p, T, mu, e_int)
template <bool DT>
__global__ __launch_bounds__(256) void k_prim(const BlockDev* __restrict__ blks,
                                              const int* __restrict__ wg_prefix, int nblk,
                                              Ptrs q, NumDev nd, double* __restrict__ part,
                                              int zfuse, SpongeHead sh) {
  if ((int)blockIdx.x < sh.head) {
    sponge_head(blks, nblk, q, sh, (int)blockIdx.x);
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
  int I = 0, J = 0, K = 0;
  if (c < cells) cell_ijk(B, (unsigned)c, &I, &J, &K);
  if (c < cells && rec_read(B, I, J, K)) {
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
    if (DT && interior_of(B, (unsigned)c, &I, &J, &K)) dtc = cell_dt(B, nd, (unsigned)c, I, J, K, w);
  }
  if (DT) {
    const double r = block_reduce<true>(dtc);
    if (threadIdx.x == 0) part[wg] = r;
  }
}

template <bool DT>
__global__ __launch_bounds__(256) void k_prim32(const BlockDev* __restrict__ blks,
                                                const int* __restrict__ wg_prefix, int nblk,
                                                Ptrs q, NumDev nd, double* __restrict__ part,
                                                int zfuse, SpongeHead sh) {
  prim32_block<DT>(blks, wg_prefix, nblk, q, nd, part, zfuse, sh, (int)blockIdx.x, (int)gridDim.x);
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
// CFL metric of every cell of a block (cell_metric: the cell means of the face vectors and
// their norms; the face arrays hold N + 1 faces along each direction), made once at
// workspace binding.  *nz collects (bitwise OR) the planes that are not +0.0 in some
// interior cell; the dt skips the loads of the others.
__global__ __launch_bounds__(256) void k_cell_sb(BlockDev B, unsigned* __restrict__ nz) {
  const unsigned cells = (unsigned)B.cells;
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned bits = 0;
  if (c < cells) {
    const unsigned N1 = (unsigned)B.N[1], N2 = (unsigned)B.N[2];
    const int K = (int)(c % N2), J = (int)((c / N2) % N1), I = (int)(c / (N1 * N2));
    double m[NSB];
    cell_metric(B, I, J, K, m);
#pragma unroll
    for (int i = 0; i < NSB_PLANES; ++i) B.Sb[(size_t)i * cells + c] = m[i];
    const int ng = B.ng;
    if (I >= ng && I < ng + B.n[0] && J >= ng && J < ng + B.n[1] && K >= ng && K < ng + B.n[2]) {
#pragma unroll
      for (int i = 0; i < NSB_PLANES; ++i)
        if (__double_as_longlong(m[i]) != 0) bits |= 1u << i;
    }
  }
  // every lane of the wavefront takes part in the OR; the atomic only for bits the word
  // does not hold yet (all wavefronts on one word would serialise at the L2)
  for (int o = 32; o > 0; o >>= 1) bits |= (unsigned)__shfl_xor((int)bits, o, 64);
  if ((threadIdx.x & 63) == 0 && (bits & ~*(volatile unsigned*)nz) != 0u) atomicOr(nz, bits);
}

hipError_t launch_cell_sb(const BlockDev& hb, unsigned* d_nz, hipStream_t s) {
  if (!hb.Sb) return hipSuccess;
  const long long nb = (hb.cells + 255) / 256;
  hipLaunchKernelGGL(k_cell_sb, dim3((unsigned)nb), dim3(256), 0, s, hb, d_nz);
  return hipGetLastError();
}

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
  if (!sh.pre) sh.nwg = 0;
  if (!sh.wall) sh.wall_wgx = sh.nwall = 0;
  sh.head = ((sh.nwg + sh.wall_wgx * sh.nwall) + 7) / 8 * 8;   // (keeps the XCD spans)
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

// FP32-bulk gradient / SGS / Ducros workgroup body (grad32_block) with its head work
// (per-case dt min, Newton wall faces), shared by the k_grad32 kernel (rhs.hip) and the
// persistent stage kernel (stage_fast.hip); the unit that includes it fixes the FP
// contraction (fast).
#pragma once
#include "wallface.h"

namespace wmx {

// block of workgroup wg: the largest b < nblk with prefix[b] <= wg (prefix non-decreasing,
// prefix[0] = 0), by bisection (an ensemble plan has up to 64 blocks)
__device__ __forceinline__ int find_block(const int* __restrict__ prefix, int nblk, int wg) {
  int lo = 0, hi = nblk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= wg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double nan_min(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}

// dt of case cs: the min of the case's CFL partials (list == nullptr: the identity list),
// one 256-thread workgroup.  Every thread issues U loads before it reduces them: the
// indices are clamped into the case's range (a repeated element leaves a min unchanged),
// so no load sits behind a branch and all U are in flight together (a channel of 4.7 M
// ghosted cells has 18 k partials: five memory round trips).
__device__ void case_dt_min(const int* __restrict__ list, const int* __restrict__ off,
                            const double* __restrict__ part, int cs, double* __restrict__ out) {
  constexpr int U = 16;
  __shared__ double red[4];
  const int lo = off[cs], hi = off[cs + 1];
  const int step = U * (int)blockDim.x;
  double a = __builtin_huge_val();
  if (hi > lo) {
    if (list) {
      for (int t0 = lo + (int)threadIdx.x; t0 < hi; t0 += step) {
        int ix[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ix[u] = ldu(list, (unsigned)min(t0 + u * (int)blockDim.x, hi - 1));
        double va[U];
#pragma unroll
        for (int u = 0; u < U; ++u) va[u] = ldu(part, (unsigned)ix[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) a = nan_min(a, va[u]);
      }
    } else {
      for (int t0 = lo + (int)threadIdx.x; t0 < hi; t0 += step) {
        double va[U];
#pragma unroll
        for (int u = 0; u < U; ++u) va[u] = ldu(part, (unsigned)min(t0 + u * (int)blockDim.x, hi - 1));
#pragma unroll
        for (int u = 0; u < U; ++u) a = nan_min(a, va[u]);
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) a = nan_min(a, __shfl_xor(a, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = nan_min(r, red[w]);
    out[cs] = r;
  }
}

// The launch's first `head` workgroups (a multiple of 8, so the gradient workgroups keep
// their XCD spans) do latency-bound side work beside the memory-bound gradients:
//   [0, ndt): the per-case dt min of the CFL partials (the step's first stage);
//   [ndt, ndt + wall_wgx * nwall_ops): the Newton wall faces, wall_wgx per wall op.
// Workgroup wgx of nwg of one gradient launch; called by every thread of a 256-thread
// workgroup (the k_grad32 kernel, and the persistent stage kernel per workgroup index).
template <int SGS>
__device__ __forceinline__ void grad32_block(const BlockDev* __restrict__ blks,
                                             const int* __restrict__ wg_prefix, int nblk,
                                             const NumDev& nd, int head,
                                             const WallOp* __restrict__ wall_ops, int wall_wgx,
                                             int nwall_ops, const Ptrs& q, int ndt,
                                             const int* __restrict__ dt_list,
                                             const int* __restrict__ dt_off,
                                             const double* __restrict__ dt_part,
                                             double* __restrict__ dt_out, int wgx, int nwg) {
  if (wgx < head) {
    const int h = wgx;
    if (h < ndt) {
      case_dt_min(dt_list, dt_off, dt_part, h, dt_out);
      return;
    }
    const int op_i = (h - ndt) / max(wall_wgx, 1);
    if (wall_wgx <= 0 || op_i >= nwall_ops) return;      // padding workgroups
    const WallOp op = wall_ops[op_i];
    const int idx0 = (h - ndt - op_i * wall_wgx) * (int)blockDim.x + (int)threadIdx.x;
    if (idx0 < op.count) wall_face<false>(blks, op, idx0, q, nd, nullptr, nullptr);
    return;
  }
  const int wg = xcd_span(wgx - head, nwg - head);
  const int b = find_block(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  const int ng = B.ng;
  const int b0 = B.n[0] + 2, b1 = B.n[1] + 2, b2 = B.n[2] + 2;
  constexpr int RUN = grad_run(sizeof(float));
  const int nrun = (b0 + RUN - 1) / RUN;
  const int l = (wg - wg_prefix[b]) * (int)blockDim.x + (int)threadIdx.x;
  if (l >= nrun * b1 * b2) return;
  const int K = ng - 1 + l % b2;
  const int J = ng - 1 + (l / b2) % b1;
  const int Ia = ng - 1 + (l / (b1 * b2)) * RUN;
  const int Ib = min(Ia + RUN, ng + B.n[0] + 1);
  const int N1 = B.N[1], N2 = B.N[2];
  const unsigned cells = (unsigned)B.cells;
  const unsigned st0 = (unsigned)N1 * N2, st1 = (unsigned)N2;
  // face-array strides: S[d] has N_d + 1 entries along d
  const unsigned f0s0 = st0;                              // i faces: (I, J, K) -> I N1 N2 + J N2 + K
  const unsigned f1s0 = (unsigned)(N1 + 1) * N2;          // j faces: M1 = N1 + 1
  const unsigned f2s0 = (unsigned)N1 * (N2 + 1), f2s1 = (unsigned)(N2 + 1);
  const unsigned fc0 = (unsigned)fcount(B.N, 0), fc1 = (unsigned)fcount(B.N, 1),
                 fc2 = (unsigned)fcount(B.N, 2);
  const float* Pf = B.Pf;
  // (u, v, w, T') of a cell: the first 16-byte group of its primitive record
  auto phi = [&](unsigned cc, float* o) {
    const float4 a = ld4(Pf + 4 * (size_t)cc);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  };
  // float copies of the face vectors and volumes (the arithmetic is FP32; same values)
  const float* S0[3] = {B.Sf[0], B.Sf[0] + fc0, B.Sf[0] + 2 * (size_t)fc0};
  const float* S1[3] = {B.Sf[1], B.Sf[1] + fc1, B.Sf[1] + 2 * (size_t)fc1};
  const float* S2[3] = {B.Sf[2], B.Sf[2] + fc2, B.Sf[2] + 2 * (size_t)fc2};
  float* Gf = B.Gf;
  const float h = 0.5f;
  const GasC& gs = nd.gas;
  unsigned c = (unsigned)Ia * st0 + (unsigned)J * st1 + (unsigned)K;
  unsigned f0 = c;                                        // i-face index of face (I, J, K)
  unsigned f1 = (unsigned)Ia * f1s0 + (unsigned)J * st1 + (unsigned)K;
  unsigned f2 = (unsigned)Ia * f2s0 + (unsigned)J * f2s1 + (unsigned)K;
  float phm[4], ph0[4], S0lo[3];
  phi(c - st0, phm);
  phi(c, ph0);
#pragma unroll
  for (int x = 0; x < 3; ++x) S0lo[x] = ldu(S0[x], f0);
#pragma unroll 1
  for (int I = Ia; I < Ib; ++I) {
    float php[4], S0hi[3];
    phi(c + st0, php);
#pragma unroll
    for (int x = 0; x < 3; ++x) S0hi[x] = ldu(S0[x], f0 + f0s0);
    float acc[12];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float flo_v = h * (phm[v] + ph0[v]);
      const float fhi_v = h * (ph0[v] + php[v]);
#pragma unroll
      for (int x = 0; x < 3; ++x) acc[3 * v + x] = fhi_v * S0hi[x] - flo_v * S0lo[x];
    }
    {
      float Slo[3], Shi[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        Slo[x] = ldu(S1[x], f1);
        Shi[x] = ldu(S1[x], f1 + st1);
      }
      float pm[4], pp[4];
      phi(c - st1, pm);
      phi(c + st1, pp);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float flo_v = h * (pm[v] + ph0[v]);
        const float fhi_v = h * (ph0[v] + pp[v]);
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * Shi[x] - flo_v * Slo[x]);
      }
    }
    {
      float Slo[3], Shi[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        Slo[x] = ldu(S2[x], f2);
        Shi[x] = ldu(S2[x], f2 + 1);
      }
      float pm[4], pp[4];
      phi(c - 1, pm);
      phi(c + 1, pp);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float flo_v = h * (pm[v] + ph0[v]);
        const float fhi_v = h * (ph0[v] + pp[v]);
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * Shi[x] - flo_v * Slo[x]);
      }
    }
    const float volr = ldu((const float*)B.volf, c);
    float g[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) g[i] = fdiv<float>(acc[i], volr);
    const float4 pb = ld4(Pf + 4 * ((size_t)cells + c));     // (rho, p', mu, e')
    const float rho = pb.x, mu = pb.z;
    float mut;
    if constexpr (SGS != WMX_SGS_NONE) {
      // filter width vol^(1/3) (solver.py:82), single-precision cube root
      const float delta = cbrtf(volr);
      const float cs = (float)nd.sgs_c;
      float nut;
      if constexpr (SGS == WMX_SGS_SMAGORINSKY) nut = nut_smagorinsky<float>(g, delta, cs);
      else if constexpr (SGS == WMX_SGS_WALE) nut = nut_wale<float>(g, delta, cs);
      else nut = nut_vreman<float>(g, delta, cs);
      mut = rho * nut;
    } else {
      mut = 0.0f * rho;
    }
    const size_t gq = 4 * (size_t)cells;                  // one float4 plane
    float* grec = Gf + 4 * (size_t)c;
    st4(grec, make_float4(g[0], g[1], g[2], g[3]));
    st4(grec + gq, make_float4(g[4], g[5], g[6], g[7]));
    st4(grec + 2 * gq, make_float4(g[8], g[9], g[10], g[11]));
    st4(grec + 3 * gq, make_float4(mu + mut, (float)gs.cp * (fdiv<float>(mu, (float)gs.Pr) + fdiv<float>(mut, (float)gs.Pr_t)),
                                   ducros<float>(g), 0.0f));
#pragma unroll
    for (int v = 0; v < 4; ++v) { phm[v] = ph0[v]; ph0[v] = php[v]; }
#pragma unroll
    for (int x = 0; x < 3; ++x) S0lo[x] = S0hi[x];
    c += st0;
    f0 += f0s0;
    f1 += f1s0;
    f2 += f2s0;
  }
}

}  // namespace wmx

// Right-hand side of the explicit step and the fused SSP-RK3 stage update.
//
//   k_grad     cells [ng-1, ng+n] per axis: Green-Gauss grad of [u v w T], SGS mu_t,
//              Ducros theta, mu_eff, k_eff                            (viscous.py:6-25,
//                                                                       sgs.py, solver.py:78-85)
//   k_wall     WallModel faces: wall flux, tau_w, q_w                 (bc.py:145-166),
//              with the Newton model or the learned model of wallnet.h
//   (the primitive kernels, q -> rho u v w p T mu e_int for all ghosted cells (gas.py:24-33),
//   are in reduce.hip with the fused CFL partial; the fused face-flux / divergence / stage
//   kernel is in flux.hip)
#include "grad32.h"

namespace wmx {


// ------------------------------------------------------------------------------ grad
// FP64: one thread per cell.
template <class R>
__global__ __launch_bounds__(256) void k_grad_cell(const BlockDev* __restrict__ blks,
                                              const int* __restrict__ wg_prefix, int nblk,
                                              NumDev nd) {
  const int wg = blockIdx.x;
  const int b = find_block(wg_prefix, nblk, wg);
  const BlockDev& B = blks[b];
  const int ng = B.ng;
  const int b0 = B.n[0] + 2, b1 = B.n[1] + 2, b2 = B.n[2] + 2;
  long long l = (long long)(wg - wg_prefix[b]) * blockDim.x + threadIdx.x;
  if (l >= (long long)b0 * b1 * b2) return;
  const int K = ng - 1 + (int)(l % b2);
  const int J = ng - 1 + (int)((l / b2) % b1);
  const int I = ng - 1 + (int)(l / ((long long)b1 * b2));
  const int* N = B.N;
  const long long cells = B.cells;
  const long long c = cidx(N, I, J, K);
  const long long st[3] = {(long long)N[1] * N[2], (long long)N[2], 1};
  const R* P = sizeof(R) == 8 ? (const R*)B.P : (const R*)B.Pf;
  const int pl[4] = {P_U, P_V, P_W, P_T};
  R phi0[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) phi0[v] = P[pl[v] * cells + c];
  R acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = R(0.0);
  int cc[3] = {I, J, K};
  const R h = R(0.5);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* S = B.S[d];
    const long long fc = fcount(N, d);
    int hi[3] = {cc[0], cc[1], cc[2]};
    hi[d] += 1;
    const long long flo = fidx(N, d, cc[0], cc[1], cc[2]);
    const long long fhi = fidx(N, d, hi[0], hi[1], hi[2]);
    R Slo[3] = {R(S[flo]), R(S[fc + flo]), R(S[2 * fc + flo])};
    R Shi[3] = {R(S[fhi]), R(S[fc + fhi]), R(S[2 * fc + fhi])};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      R pm = P[pl[v] * cells + c - st[d]];
      R pp = P[pl[v] * cells + c + st[d]];
      R flo_v = h * (pm + phi0[v]);
      R fhi_v = h * (phi0[v] + pp);
#pragma unroll
      for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * Shi[x] - flo_v * Slo[x]);
    }
  }
  const double vol = B.vol[c];
  const R volr = R(vol);
  // one reciprocal of the volume (within an ulp of the reference's twelve divisions)
  const R ivol = fdiv<R>(R(1.0), volr);
  R g[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) g[i] = acc[i] * ivol;
  const R rho = P[c];
  const R mu = P[P_MU * cells + c];
  R mut = R(0.0);
  if (nd.sgs != WMX_SGS_NONE) {
    // filter width vol^(1/3) (solver.py:82); FP32-bulk: single-precision cube root
    R delta;
    if constexpr (sizeof(R) == 4) delta = cbrtf((float)vol);
    else delta = R(cbrt(vol));                   // vol ** (1/3), within an ulp
    const R cs = R(nd.sgs_c);
    R nut;
    if (nd.sgs == WMX_SGS_SMAGORINSKY) nut = nut_smagorinsky<R>(g, delta, cs);
    else if (nd.sgs == WMX_SGS_WALE) nut = nut_wale<R>(g, delta, cs);
    else nut = nut_vreman<R>(g, delta, cs);
    mut = rho * nut;
  } else {
    mut = R(0.0) * rho;
  }
  R* G = sizeof(R) == 8 ? (R*)B.G : (R*)B.Gf;
#pragma unroll
  for (int i = 0; i < 12; ++i) G[i * cells + c] = g[i];
  const GasC& gs = nd.gas;
  G[G_MU * cells + c] = mu + mut;
  G[G_K * cells + c] = R(gs.cp) * (fdiv<R>(mu, R(gs.Pr)) + fdiv<R>(mut, R(gs.Pr_t)));
  G[G_TH * cells + c] = ducros<R>(g);
}

template <int SGS>
__global__ __launch_bounds__(256) void k_grad32(const BlockDev* __restrict__ blks,
                                                const int* __restrict__ wg_prefix, int nblk,
                                                NumDev nd, int head, const WallOp* __restrict__ wall_ops,
                                                int wall_wgx, int nwall_ops, Ptrs q, int ndt,
                                                const int* __restrict__ dt_list,
                                                const int* __restrict__ dt_off,
                                                const double* __restrict__ dt_part,
                                                double* __restrict__ dt_out) {
  grad32_block<SGS>(blks, wg_prefix, nblk, nd, head, wall_ops, wall_wgx, nwall_ops, q, ndt, dt_list, dt_off,
                    dt_part, dt_out, (int)blockIdx.x, (int)gridDim.x);
}
template __global__ void k_grad32<WMX_SGS_NONE>(const BlockDev*, const int*, int, NumDev, int,
                                                  const WallOp*, int, int, Ptrs, int, const int*,
                                                  const int*, const double*, double*);
template __global__ void k_grad32<WMX_SGS_SMAGORINSKY>(const BlockDev*, const int*, int, NumDev, int,
                                                  const WallOp*, int, int, Ptrs, int, const int*,
                                                  const int*, const double*, double*);
template __global__ void k_grad32<WMX_SGS_WALE>(const BlockDev*, const int*, int, NumDev, int,
                                                  const WallOp*, int, int, Ptrs, int, const int*,
                                                  const int*, const double*, double*);
template __global__ void k_grad32<WMX_SGS_VREMAN>(const BlockDev*, const int*, int, NumDev, int,
                                                  const WallOp*, int, int, Ptrs, int, const int*,
                                                  const int*, const double*, double*);

// ------------------------------------------------------------------------------ wall
// NET: the learned model (wave-cooperative network + Newton polish) replaces the
// Newton solve; whole waves past the face count leave, the others clamp their index.
template <bool NET>
__global__ __launch_bounds__(256) void k_wall(const BlockDev* __restrict__ blks,
                                              const WallOp* __restrict__ ops, Ptrs q, NumDev nd) {
  __shared__ unsigned short sW[NET ? MLP_LDS_BF16 : 1];
  __shared__ float sF[NET ? MLP_NF32 : 1];
  const WallOp op = ops[blockIdx.y];
  const int idx0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (NET) {
    net_stage(nd.net, sW, sF);
    __syncthreads();
    if ((int)(blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >= op.count) return;
  } else if (idx0 >= op.count) {
    return;
  }
  wall_face<NET>(blks, op, idx0, q, nd, sW, sF);
}

// ------------------------------------------------------------------- shells / sponge
// Both shell kernels take every load of a cell before its first store: a store between
// two loads would make each of the five components a separate memory round trip (the
// compiler cannot move a load across a store that may alias it).
// save: shell <- q ghosts
__global__ void k_shell_save(const BlockDev* __restrict__ blks, const int* __restrict__ wg_prefix,
                             int nblk, Ptrs q) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  const unsigned s = (blockIdx.x - (unsigned)wg_prefix[b]) * blockDim.x + threadIdx.x;
  const size_t nsh = (size_t)B.shell_cells, cells = (size_t)B.cells;
  if (s >= nsh) return;
  int I, J, K;
  shell_cell(s, B.N, B.n, B.ng, &I, &J, &K);
  // (interior-(I, J) k ghosts of fused sides: q0 keeps them through the step)
  if (I >= B.ng && I < B.ng + B.n[0] && J >= B.ng && J < B.ng + B.n[1] &&
      shell_kept(B.n, B.ng, B.zfuse, K))
    return;
  const double* src = q.p[b] + cidx(B.N, I, J, K);
  double* dst = B.shell + s;
  double g[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) g[v] = src[v * cells];
#pragma unroll
  for (int v = 0; v < 5; ++v) dst[v * nsh] = g[v];
}

// q0 ghosts <- shell (restore); out ghosts <- the three stage combinations applied to
// the saved ghosts with R = 0 there (solver.py:105,110-121); zero: out ghosts <- 0.
template <int MODE>  // 0: zero out; 1: restore q0 + RK3 ghost combination into out
__global__ void k_shell_finish(const BlockDev* __restrict__ blks, const int* __restrict__ wg_prefix,
                               int nblk, Ptrs q0, Ptrs out, const double* __restrict__ d_dt,
                               double dt_host) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  const unsigned s = (blockIdx.x - (unsigned)wg_prefix[b]) * blockDim.x + threadIdx.x;
  const size_t nsh = (size_t)B.shell_cells, cells = (size_t)B.cells;
  if (s >= nsh) return;
  int I, J, K;
  shell_cell(s, B.N, B.n, B.ng, &I, &J, &K);
  const long long c = cidx(B.N, I, J, K);
  double* oc = out.p[b] + c;
  if (MODE == 0) {
#pragma unroll
    for (int v = 0; v < 5; ++v) oc[v * cells] = 0.0;
    return;
  }
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  const double zr = dt * 0.0;
  // interior-(I, J) k ghosts of fused sides were never written in q0: read there, no restore
  const bool kept = I >= B.ng && I < B.ng + B.n[0] && J >= B.ng && J < B.ng + B.n[1] &&
                    shell_kept(B.n, B.ng, B.zfuse, K);
  double* qc = q0.p[b] + c;
  double g[5];
  if (kept) {
#pragma unroll
    for (int v = 0; v < 5; ++v) g[v] = qc[v * cells];
  } else {
    const double* sc = B.shell + s;
#pragma unroll
    for (int v = 0; v < 5; ++v) g[v] = sc[v * nsh];
#pragma unroll
    for (int v = 0; v < 5; ++v) qc[v * cells] = g[v];
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double q1 = g[v] + zr;
    const double q2 = 0.75 * g[v] + 0.25 * (q1 + zr);
    oc[v * cells] = (1.0 / 3.0) * g[v] + (2.0 / 3.0) * (q2 + zr);
  }
}

template __global__ void k_shell_finish<0>(const BlockDev*, const int*, int, Ptrs, Ptrs,
                                           const double*, double);
template __global__ void k_shell_finish<1>(const BlockDev*, const int*, int, Ptrs, Ptrs,
                                           const double*, double);

// Wall model on point arrays (utau / __call__ of the reference model object).
__global__ void k_wm_eval(NumDev nd, int mode, long long n, const double* __restrict__ a,
                          const double* __restrict__ b, const double* __restrict__ c,
                          const double* __restrict__ d, double T_wall, double* __restrict__ o0,
                          double* __restrict__ o1) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mode == 0) {
    o0[i] = spalding_utau(nd.wm, a[i], b[i], c[i], d[i]);
  } else {
    double tau, qw;
    equilibrium_wall(nd.gas, nd.wm, a[i], b[i], c[i], d[i], T_wall, &tau, &qw);
    o0[i] = tau;
    o1[i] = qw;
  }
}

// float copy of a double array (FP32-bulk geometry copies)
__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ----------------------------------------------------------------------- launchers
hipError_t launch_to_f32(const double* a, float* b, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long nb = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, s, a, b, n);
  return hipGetLastError();
}

hipError_t launch_wm_eval(const NumDev& nd, int mode, long long n, const double* a, const double* b,
                          const double* c, const double* d, double T_wall, double* o0, double* o1,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wm_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nd, mode, n, a,
                     b, c, d, T_wall, o0, o1);
  return hipGetLastError();
}

hipError_t launch_grad(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, const NumDev& nd,
                       const WallOp* d_wall, int nwall_ops, int wall_max, Ptrs q, const GradDt& gdt,
                       hipStream_t s) {
  if (nd.prec == WMX_PREC_FP32_BULK) {
    // side work in the same launch: per-case dt mins, then Newton wall faces (padded to 8);
    // nwg == 0 (fused plans): the side work alone
    const int wx = d_wall && nwall_ops > 0 && wall_max > 0 ? (wall_max + 255) / 256 : 0;
    const int ndt = gdt.out ? gdt.ncases : 0;
    const int head = ((ndt + wx * nwall_ops + 7) / 8) * 8;
    if (nwg + head <= 0) return hipSuccess;
    const dim3 grid(nwg + head);
#define WMX_GRAD32(S)                                                                          \
  hipLaunchKernelGGL(k_grad32<S>, grid, dim3(256), 0, s, d_blks, d_pre, nblk, nd, head, d_wall, \
                     wx, nwall_ops, q, ndt, gdt.list, gdt.off, gdt.part, gdt.out)
    switch (nd.sgs) {
      case WMX_SGS_NONE: WMX_GRAD32(WMX_SGS_NONE); break;
      case WMX_SGS_SMAGORINSKY: WMX_GRAD32(WMX_SGS_SMAGORINSKY); break;
      case WMX_SGS_WALE: WMX_GRAD32(WMX_SGS_WALE); break;
      default: WMX_GRAD32(WMX_SGS_VREMAN);
    }
#undef WMX_GRAD32
  } else if (nwg > 0)
    hipLaunchKernelGGL(k_grad_cell<double>, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, nd);
  return hipGetLastError();
}

hipError_t launch_wall(const BlockDev* d_blks, const WallOp* d_ops, int nops, int maxcount, Ptrs q,
                       const NumDev& nd, hipStream_t s) {
  if (nops <= 0 || maxcount <= 0) return hipSuccess;
  dim3 grid((maxcount + 255) / 256, nops);
  if (nd.net.w)
    hipLaunchKernelGGL(k_wall<true>, grid, dim3(256), 0, s, d_blks, d_ops, q, nd);
  else
    hipLaunchKernelGGL(k_wall<false>, grid, dim3(256), 0, s, d_blks, d_ops, q, nd);
  return hipGetLastError();
}

hipError_t launch_shell_save(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_shell_save, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q);
  return hipGetLastError();
}

hipError_t launch_shell_finish(int mode, const BlockDev* d_blks, const int* d_pre, int nblk, int nwg,
                               Ptrs q0, Ptrs out, const double* d_dt, double dt, hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  if (mode == 0)
    hipLaunchKernelGGL(k_shell_finish<0>, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q0, out, d_dt, dt);
  else
    hipLaunchKernelGGL(k_shell_finish<1>, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q0, out, d_dt, dt);
  return hipGetLastError();
}

}  // namespace wmx

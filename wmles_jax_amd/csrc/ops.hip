// Operator-level entry points of the reference's plugin interfaces, evaluated by the
// same point-wise device code as the fused step (physics.h), in FP64:
//
//   k_scheme_flux   scheme(blk, geom, w, gas, d, theta) -> F: the inviscid flux (incl.
//                   area) of KEP / KEEP / MUSCL-HLLC / Hybrid through the n_d + 1 interior
//                   faces of direction d over the interior tangential cells
//                   (reference wmles/schemes.py:3-5,57-166)
//   k_sgs_eval      make_sgs(name)(g, delta, rho) -> mu_t (reference wmles/sgs.py:45-50)
//   k_wall_flux     WallModel.wall_flux(w, T, blk, geom, gas, axis, side) -> (F, tau, q_w)
//                   with the equilibrium model (reference wmles/bc.py:145-166)
#include "common.h"

namespace wmx {

namespace {

__device__ __forceinline__ CellW cell_w(const double* w, long long cells, long long c, double gm1) {
  CellW o;
  o.r = w[c];
  o.u = w[cells + c];
  o.v = w[2 * cells + c];
  o.w = w[3 * cells + c];
  o.p = w[4 * cells + c];
  o.e = o.p / (gm1 * o.r);          // as the primitive kernel stores e_int
  return o;
}

// One thread per face; faces in C order of the output (n_d + 1 along d).
__global__ __launch_bounds__(256) void k_scheme_flux(NumDev nd, int3 n, int d, const double* __restrict__ w,
                                                     const double* __restrict__ S,
                                                     const double* __restrict__ theta,
                                                     double* __restrict__ F) {
  const int ng = 3;
  const int m0 = n.x + (d == 0), m1 = n.y + (d == 1), m2 = n.z + (d == 2);
  const long long nf = (long long)m0 * m1 * m2;
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int a2 = (int)(f % m2), a1 = (int)((f / m2) % m1), a0 = (int)(f / ((long long)m1 * m2));
  const int N[3] = {n.x + 2 * ng, n.y + 2 * ng, n.z + 2 * ng};
  const long long cells = (long long)N[0] * N[1] * N[2];
  const int I = ng + a0, J = ng + a1, K = ng + a2;             // right cell of the face
  const long long sd = d == 0 ? (long long)N[1] * N[2] : (d == 1 ? (long long)N[2] : 1LL);
  const long long cR = cidx(N, I, J, K), cL = cR - sd;
  const long long fi = fidx(N, d, I, J, K), fc = fcount(N, d);
  const double Sx = S[fi], Sy = S[fc + fi], Sz = S[2 * fc + fi];
  const double gm1 = nd.gas.gamma - 1.0;
  const CellW L = cell_w(w, cells, cL, gm1), R = cell_w(w, cells, cR, gm1);
  const int sch = nd.scheme;
  double out[5];
  const bool keep_c = nd.central == WMX_SCHEME_KEEP;
  if (sch == WMX_SCHEME_KEP || sch == WMX_SCHEME_KEEP) {
    CellW L2 = L, R2 = R;
    if (nd.order == 4) { L2 = cell_w(w, cells, cL - sd, gm1); R2 = cell_w(w, cells, cR + sd, gm1); }
    if (sch == WMX_SCHEME_KEP) central_flux<false>(nd.order, L2, L, R, R2, Sx, Sy, Sz, out);
    else central_flux<true>(nd.order, L2, L, R, R2, Sx, Sy, Sz, out);
  } else {
    const CellW L2 = cell_w(w, cells, cL - sd, gm1), R2 = cell_w(w, cells, cR + sd, gm1);
    if (sch == WMX_SCHEME_MUSCLHLLC) {
      muscl_hllc_flux(nd.gas.gamma, nd.limiter, L2, L, R, R2, Sx, Sy, Sz, out);
    } else {
      // Hybrid (schemes.py:162-166); a part of weight exactly 0 is not evaluated, which
      // is the same value while the skipped flux is finite
      const double th = fmax(theta[cL], theta[cR]);
      const double s = fmin(fmax((th - nd.th0) / (nd.th1 - nd.th0), 0.0), 1.0);
      double Fc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Fu[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (s < 1.0) {
        CellW L2 = L, R2 = R;
        if (nd.order == 4) { L2 = cell_w(w, cells, cL - sd, gm1); R2 = cell_w(w, cells, cR + sd, gm1); }
        if (keep_c) central_flux<true>(nd.order, L2, L, R, R2, Sx, Sy, Sz, Fc);
        else central_flux<false>(nd.order, L2, L, R, R2, Sx, Sy, Sz, Fc);
      }
      if (s > 0.0) {
        const CellW L2 = cell_w(w, cells, cL - sd, gm1), R2 = cell_w(w, cells, cR + sd, gm1);
        muscl_hllc_flux(nd.gas.gamma, nd.limiter, L2, L, R, R2, Sx, Sy, Sz, Fu);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] = (1.0 - s) * Fc[v] + s * Fu[v];
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) F[v * nf + f] = out[v];
}

// g: (3, 3, n) velocity gradient g[i][j] = du_i/dx_j; delta, rho: (n,)
__global__ __launch_bounds__(256) void k_sgs_eval(NumDev nd, long long n, const double* __restrict__ g,
                                                  const double* __restrict__ delta,
                                                  const double* __restrict__ rho, double* __restrict__ mut) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double gv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) gv[k] = g[k * n + c];
  const double r = rho[c];
  if (nd.sgs == WMX_SGS_NONE) {
    mut[c] = 0.0 * r;
    return;
  }
  const double dl = delta[c], cs = nd.sgs_c;
  double nut;
  if (nd.sgs == WMX_SGS_SMAGORINSKY) nut = nut_smagorinsky<double>(gv, dl, cs);
  else if (nd.sgs == WMX_SGS_WALE) nut = nut_wale<double>(gv, dl, cs);
  else nut = nut_vreman<double>(gv, dl, cs);
  mut[c] = r * nut;
}

// One thread per wall face (t1, t2) of face (a, s); w: (5, N...) primitives, T: (N...).
__global__ __launch_bounds__(256) void k_wall_flux(NumDev nd, BlockDev B, int a, int s,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ T, double* __restrict__ F,
                                                   double* __restrict__ tau_out, double* __restrict__ qw_out) {
  int t1, t2;
  tangents(a, &t1, &t2);
  const long long nw = (long long)B.n[t1] * B.n[t2];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nw) return;
  const int ng = B.ng, n = B.n[a];
  const wmx_face_desc& fd = B.faces[2 * a + s];
  int c[3];
  c[t2] = ng + (int)(idx % B.n[t2]);
  c[t1] = ng + (int)(idx / B.n[t2]);
  const int fb = s == 0 ? ng : ng + n;
  const int mi = s == 0 ? ng + fd.n_match - 1 : ng + n - fd.n_match;
  const int ci = s == 0 ? ng : ng + n - 1;
  const int* N = B.N;
  const long long cells = B.cells;
  int cf[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]}, cb[3] = {c[0], c[1], c[2]};
  cf[a] = fb; cm[a] = mi; cb[a] = ci;
  const long long fi = fidx(N, a, cf[0], cf[1], cf[2]), fc = fcount(N, a);
  const double* S = B.S[a];
  const double* xf = B.xf[a];
  const double Sv[3] = {S[fi], S[fc + fi], S[2 * fc + fi]};
  const double Sm = sqrt((Sv[0] * Sv[0] + Sv[1] * Sv[1]) + Sv[2] * Sv[2]);
  double no[3];
  for (int x = 0; x < 3; ++x) no[x] = s == 1 ? Sv[x] / Sm : -Sv[x] / Sm;
  const long long m = cidx(N, cm[0], cm[1], cm[2]);
  const long long cbi = cidx(N, cb[0], cb[1], cb[2]);
  const double h = (((B.xc[m] - xf[fi]) * (-no[0]) + (B.xc[cells + m] - xf[fc + fi]) * (-no[1])) +
                    (B.xc[2 * cells + m] - xf[2 * fc + fi]) * (-no[2]));
  const double um[3] = {w[cells + m], w[2 * cells + m], w[3 * cells + m]};
  const double un = dot3(um[0], um[1], um[2], no[0], no[1], no[2]);
  const double up[3] = {um[0] - un * no[0], um[1] - un * no[1], um[2] - un * no[2]};
  const double U = sqrt(((up[0] * up[0] + up[1] * up[1]) + up[2] * up[2]) + 1e-20);
  double tau, qw;
  equilibrium_wall(nd.gas, nd.wm, h, U, T[m], w[4 * cells + m], fd.T_wall, &tau, &qw);
  const double pb = w[4 * cells + cbi];
  const double sg = s == 1 ? 1.0 : -1.0;
  F[idx] = sg * 0.0;
  for (int x = 0; x < 3; ++x) F[(1 + x) * nw + idx] = sg * ((pb * no[x] + tau * (up[x] / U)) * Sm);
  F[4 * nw + idx] = sg * ((-qw) * Sm);
  tau_out[idx] = tau;
  qw_out[idx] = qw;
}

}  // namespace

hipError_t launch_scheme_flux(const NumDev& nd, const int* n, int d, const double* w, const double* S,
                              const double* theta, double* F, hipStream_t s) {
  const long long nf = (long long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scheme_flux, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, nd,
                     make_int3(n[0], n[1], n[2]), d, w, S, theta, F);
  return hipGetLastError();
}

hipError_t launch_sgs_eval(const NumDev& nd, long long n, const double* g, const double* delta,
                           const double* rho, double* mut, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sgs_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nd, n, g, delta, rho, mut);
  return hipGetLastError();
}

hipError_t launch_wall_flux(const NumDev& nd, const BlockDev& B, int a, int side, const double* w,
                            const double* T, double* F, double* tau, double* qw, hipStream_t s) {
  int t1, t2;
  tangents(a, &t1, &t2);
  const long long nw = (long long)B.n[t1] * B.n[t2];
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wall_flux, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, nd, B, a, side, w, T,
                     F, tau, qw);
  return hipGetLastError();
}

}  // namespace wmx

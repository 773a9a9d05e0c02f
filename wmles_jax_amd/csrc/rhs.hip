// Right-hand side of the explicit step and the fused SSP-RK3 stage update.
//
//   k_prim     all ghosted cells: q -> rho u v w p T mu e_int          (gas.py:24-33)
//   k_grad     cells [ng-1, ng+n] per axis: Green-Gauss grad of [u v w T], SGS mu_t,
//              Ducros theta, mu_eff, k_eff                            (viscous.py:6-25,
//                                                                       sgs.py, solver.py:78-85)
//   k_wall     WallModel faces: wall flux, tau_w, q_w                 (bc.py:145-166)
//   k_flux     2.5-D march: (j,k) tile of 16x16 cells walking i; all face fluxes
//              (inviscid + viscous, wall override) once per face, divergence / vol,
//              sources, RK stage combination                          (solver.py:86-95,
//                                                                       103-121)
#include "common.h"

namespace wmx {

constexpr int TJ = 16, TK = 16;

__device__ __forceinline__ int find_block(const int* __restrict__ prefix, int nblk, int wg) {
  int b = 0;
  while (b + 1 < nblk && prefix[b + 1] <= wg) ++b;
  return b;
}

// ------------------------------------------------------------------------------ prim
__global__ __launch_bounds__(256) void k_prim(const BlockDev* __restrict__ blks,
                                              const int* __restrict__ wg_prefix, int nblk,
                                              Ptrs q, GasC gas) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  long long c = (long long)(blockIdx.x - wg_prefix[b]) * blockDim.x + threadIdx.x;
  const long long cells = B.cells;
  if (c >= cells) return;
  const double* qb = q.p[b];
  Prim w = cons_to_prim(gas, qb[c], qb[cells + c], qb[2 * cells + c], qb[3 * cells + c],
                        qb[4 * cells + c]);
  double* P = B.P;
  P[c] = w.r;
  P[cells + c] = w.u;
  P[2 * cells + c] = w.v;
  P[3 * cells + c] = w.w;
  P[4 * cells + c] = w.p;
  P[5 * cells + c] = w.T;
  P[6 * cells + c] = mu_suth(gas, w.T);
  P[7 * cells + c] = w.p / ((gas.gamma - 1.0) * w.r);
}

// ------------------------------------------------------------------------------ grad
__global__ __launch_bounds__(256) void k_grad(const BlockDev* __restrict__ blks,
                                              const int* __restrict__ wg_prefix, int nblk,
                                              NumDev nd) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  const int ng = B.ng;
  const int b0 = B.n[0] + 2, b1 = B.n[1] + 2, b2 = B.n[2] + 2;
  long long l = (long long)(blockIdx.x - wg_prefix[b]) * blockDim.x + threadIdx.x;
  if (l >= (long long)b0 * b1 * b2) return;
  const int K = ng - 1 + (int)(l % b2);
  const int J = ng - 1 + (int)((l / b2) % b1);
  const int I = ng - 1 + (int)(l / ((long long)b1 * b2));
  const int* N = B.N;
  const long long cells = B.cells;
  const long long c = cidx(N, I, J, K);
  const long long st[3] = {(long long)N[1] * N[2], (long long)N[2], 1};
  const double* P = B.P;
  const int pl[4] = {P_U, P_V, P_W, P_T};
  double phi0[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) phi0[v] = P[pl[v] * cells + c];
  double acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = 0.0;
  int cc[3] = {I, J, K};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double* S = B.S[d];
    const long long fc = fcount(N, d);
    int hi[3] = {cc[0], cc[1], cc[2]};
    hi[d] += 1;
    const long long flo = fidx(N, d, cc[0], cc[1], cc[2]);
    const long long fhi = fidx(N, d, hi[0], hi[1], hi[2]);
    double Slo[3] = {S[flo], S[fc + flo], S[2 * fc + flo]};
    double Shi[3] = {S[fhi], S[fc + fhi], S[2 * fc + fhi]};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double pm = P[pl[v] * cells + c - st[d]];
      double pp = P[pl[v] * cells + c + st[d]];
      double flo_v = 0.5 * (pm + phi0[v]);
      double fhi_v = 0.5 * (phi0[v] + pp);
#pragma unroll
      for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * Shi[x] - flo_v * Slo[x]);
    }
  }
  const double vol = B.vol[c];
  double g[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) g[i] = acc[i] / vol;
  const double rho = P[c];
  const double mu = P[P_MU * cells + c];
  double mut = 0.0;
  if (nd.sgs != WMX_SGS_NONE) {
    const double delta = pow(vol, 1.0 / 3.0);
    double nut;
    if (nd.sgs == WMX_SGS_SMAGORINSKY) nut = nut_smagorinsky(g, delta, nd.sgs_c);
    else if (nd.sgs == WMX_SGS_WALE) nut = nut_wale(g, delta, nd.sgs_c);
    else nut = nut_vreman(g, delta, nd.sgs_c);
    mut = rho * nut;
  } else {
    mut = 0.0 * rho;
  }
  double* G = B.G;
#pragma unroll
  for (int i = 0; i < 12; ++i) G[i * cells + c] = g[i];
  const GasC& gs = nd.gas;
  G[G_MU * cells + c] = mu + mut;
  G[G_K * cells + c] = gs.cp * (mu / gs.Pr + mut / gs.Pr_t);
  G[G_TH * cells + c] = ducros(g);
}

// ------------------------------------------------------------------------------ wall
__global__ __launch_bounds__(256) void k_wall(const BlockDev* __restrict__ blks,
                                              const WallOp* __restrict__ ops, Ptrs q, NumDev nd) {
  const WallOp op = ops[blockIdx.y];
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= op.count) return;
  const BlockDev& B = blks[op.blk];
  const int a = op.axis, s = op.side, ng = B.ng, n = B.n[a];
  const wmx_face_desc& f = B.faces[2 * a + s];
  int t1, t2;
  tangents(a, &t1, &t2);
  int c[3];
  c[t2] = ng + idx % B.n[t2];
  c[t1] = ng + idx / B.n[t2];
  const int fb = s == 0 ? ng : ng + n;
  const int mi = s == 0 ? ng + f.n_match - 1 : ng + n - f.n_match;
  const int ci = s == 0 ? ng : ng + n - 1;
  const int* N = B.N;
  const long long cells = B.cells;
  int cf[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]}, cb[3] = {c[0], c[1], c[2]};
  cf[a] = fb; cm[a] = mi; cb[a] = ci;
  const long long fi = fidx(N, a, cf[0], cf[1], cf[2]), fc = fcount(N, a);
  const double* S = B.S[a];
  const double* xf = B.xf[a];
  double Sv[3] = {S[fi], S[fc + fi], S[2 * fc + fi]};
  double Sm = sqrt((Sv[0] * Sv[0] + Sv[1] * Sv[1]) + Sv[2] * Sv[2]);
  double no[3];
  for (int x = 0; x < 3; ++x) no[x] = s == 1 ? Sv[x] / Sm : -Sv[x] / Sm;
  const long long m = cidx(N, cm[0], cm[1], cm[2]);
  const long long cbi = cidx(N, cb[0], cb[1], cb[2]);
  double h = 0.0;
  {
    double d0 = (B.xc[m] - xf[fi]) * (-no[0]);
    double d1 = (B.xc[cells + m] - xf[fc + fi]) * (-no[1]);
    double d2 = (B.xc[2 * cells + m] - xf[2 * fc + fi]) * (-no[2]);
    h = (d0 + d1) + d2;
  }
  const double* qb = q.p[op.blk];
  const GasC& gas = nd.gas;
  Prim wm = cons_to_prim(gas, qb[m], qb[cells + m], qb[2 * cells + m], qb[3 * cells + m],
                         qb[4 * cells + m]);
  Prim wb = cons_to_prim(gas, qb[cbi], qb[cells + cbi], qb[2 * cells + cbi], qb[3 * cells + cbi],
                         qb[4 * cells + cbi]);
  double un = dot3(wm.u, wm.v, wm.w, no[0], no[1], no[2]);
  double up[3] = {wm.u - un * no[0], wm.v - un * no[1], wm.w - un * no[2]};
  double U = sqrt(((up[0] * up[0] + up[1] * up[1]) + up[2] * up[2]) + 1e-20);
  double tau, qw;
  equilibrium_wall(gas, nd.wm, h, U, wm.T, wm.p, f.T_wall, &tau, &qw);
  double sg = s == 1 ? 1.0 : -1.0;
  double* W = B.wall[2 * a + s];
  const long long nw = (long long)B.n[t1] * B.n[t2];
  W[idx] = sg * 0.0;
  for (int x = 0; x < 3; ++x) W[(1 + x) * nw + idx] = sg * ((wb.p * no[x] + tau * (up[x] / U)) * Sm);
  W[4 * nw + idx] = sg * ((-qw) * Sm);
  W[5 * nw + idx] = tau;
  W[6 * nw + idx] = qw;
}

// ------------------------------------------------------------------------------ flux
__device__ __forceinline__ CellW load_w(const double* __restrict__ P, long long cells, long long c) {
  CellW w;
  w.r = P[c];
  w.u = P[cells + c];
  w.v = P[2 * cells + c];
  w.w = P[3 * cells + c];
  w.p = P[4 * cells + c];
  w.e = P[7 * cells + c];
  return w;
}

__device__ __forceinline__ CellG load_g(const double* __restrict__ G, long long cells, long long c) {
  CellG g;
#pragma unroll
  for (int i = 0; i < 12; ++i) g.g[i] = G[i * cells + c];
  g.mu = G[G_MU * cells + c];
  g.k = G[G_K * cells + c];
  return g;
}

// Numerical flux (times area) through face f of direction d, f = coordinate of the
// right cell (I, J, K); solver.py:88-93 with the wall override.
__device__ __forceinline__ void face_flux(const NumDev& nd, const BlockDev& B, int d, int I, int J,
                                          int K, double* F) {
  const int ng = B.ng;
  const int cc[3] = {I, J, K};
  const int f = cc[d];
  int side = -1;
  if (f == ng) side = 0;
  else if (f == ng + B.n[d]) side = 1;
  if (side >= 0 && B.faces[2 * d + side].kind == WMX_BC_WALLMODEL) {
    int t1, t2;
    tangents(d, &t1, &t2);
    const long long nw = (long long)B.n[t1] * B.n[t2];
    const long long w = (long long)(cc[t1] - ng) * B.n[t2] + (cc[t2] - ng);
    const double* W = B.wall[2 * d + side];
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = W[v * nw + w];
    return;
  }
  const int* N = B.N;
  const long long cells = B.cells;
  const long long sd = d == 0 ? (long long)N[1] * N[2] : (d == 1 ? (long long)N[2] : 1LL);
  const long long cR = cidx(N, I, J, K), cL = cR - sd;
  const double* P = B.P;
  const double* G = B.G;
  const long long fi = fidx(N, d, I, J, K), fc = fcount(N, d);
  const double* S = B.S[d];
  const double Sx = S[fi], Sy = S[fc + fi], Sz = S[2 * fc + fi];
  const CellW L = load_w(P, cells, cL), R = load_w(P, cells, cR);
  const int sch = nd.scheme;
  if (sch == WMX_SCHEME_KEP || sch == WMX_SCHEME_KEEP) {
    CellW L2 = L, R2 = R;
    if (nd.order == 4) { L2 = load_w(P, cells, cL - sd); R2 = load_w(P, cells, cR + sd); }
    if (sch == WMX_SCHEME_KEP) central_flux<false>(nd.order, L2, L, R, R2, Sx, Sy, Sz, F);
    else central_flux<true>(nd.order, L2, L, R, R2, Sx, Sy, Sz, F);
  } else {
    const CellW L2 = load_w(P, cells, cL - sd), R2 = load_w(P, cells, cR + sd);
    if (sch == WMX_SCHEME_MUSCLHLLC) {
      muscl_hllc_flux(nd.gas.gamma, nd.limiter, L2, L, R, R2, Sx, Sy, Sz, F);
    } else {
      // Hybrid: (1 - s) F_c + s F_u; a part whose weight is exactly 0 is not
      // evaluated (bit-identical while the skipped flux is finite).
      const double th = fmax(G[G_TH * cells + cL], G[G_TH * cells + cR]);
      const double s = fmin(fmax((th - nd.th0) / (nd.th1 - nd.th0), 0.0), 1.0);
      double Fc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Fu[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (s < 1.0) {
        if (nd.central == WMX_SCHEME_KEEP) central_flux<true>(nd.order, L2, L, R, R2, Sx, Sy, Sz, Fc);
        else central_flux<false>(nd.order, L2, L, R, R2, Sx, Sy, Sz, Fc);
      }
      if (s > 0.0) muscl_hllc_flux(nd.gas.gamma, nd.limiter, L2, L, R, R2, Sx, Sy, Sz, Fu);
#pragma unroll
      for (int v = 0; v < 5; ++v) F[v] = (1.0 - s) * Fc[v] + s * Fu[v];
    }
  }
  // viscous part (viscous.py:28-51)
  const CellG gL = load_g(G, cells, cL), gR = load_g(G, cells, cR);
  const double phL[4] = {L.u, L.v, L.w, P[P_T * cells + cL]};
  const double phR[4] = {R.u, R.v, R.w, P[P_T * cells + cR]};
  const double xL[3] = {B.xc[cL], B.xc[cells + cL], B.xc[2 * cells + cL]};
  const double xR[3] = {B.xc[cR], B.xc[cells + cR], B.xc[2 * cells + cR]};
  viscous_face(gL, gR, phL, phR, xL, xR, Sx, Sy, Sz, F);
}

// MODE 0: r = RHS; 1: q + dt r; 2: 3/4 q0 + 1/4 (q + dt r); 3: 1/3 q0 + 2/3 (q + dt r)
template <int MODE>
__global__ __launch_bounds__(256) void k_flux(const BlockDev* __restrict__ blks,
                                              const int4* __restrict__ tiles, int ci_len,
                                              NumDev nd, Ptrs qin, Ptrs q0, Ptrs out,
                                              const double* __restrict__ d_dt, double dt_host) {
  __shared__ double sFj[5][(TJ + 1) * TK];
  __shared__ double sFk[5][TJ * (TK + 1)];
  const int4 t = tiles[blockIdx.x];
  const BlockDev& B = blks[t.x];
  const int j0 = t.y, k0 = t.z, i0 = t.w;
  const int i1 = min(i0 + ci_len, B.n[0]);
  const int tid = threadIdx.x;
  const int tj = tid / TK, tk = tid % TK;
  const int j = j0 + tj, k = k0 + tk;
  const bool act = j < B.n[1] && k < B.n[2];
  const int jl = min(j0 + TJ, B.n[1]) - j0;
  const int kl = min(k0 + TK, B.n[2]) - k0;
  const int ng = B.ng;
  const int J = ng + j, K = ng + k;
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  const long long cells = B.cells;
  double Flo[5], Fhi[5];
  for (int i = i0; i < i1; ++i) {
    const int I = ng + i;
    for (int task = (i == i0 ? 0 : 1); task < 5; ++task) {
      if (task == 4 && tid >= 64) break;                 // tile-edge faces: wave 0 only
      int d = 0, fI = I, fJ = J, fK = K;
      bool doit = act;
      if (task == 0) { d = 0; }
      else if (task == 1) { d = 0; fI = I + 1; }
      else if (task == 2) { d = 1; }
      else if (task == 3) { d = 2; }
      else {
        if (tid < TK) { d = 1; fJ = ng + j0 + jl; fK = ng + k0 + tid; doit = tid < kl; }
        else if (tid < TK + TJ) { d = 2; fJ = ng + j0 + (tid - TK); fK = ng + k0 + kl; doit = (tid - TK) < jl; }
        else doit = false;
      }
      double F[5];
      if (doit) face_flux(nd, B, d, fI, fJ, fK, F);
      if (doit) {
        if (task == 0) {
#pragma unroll
          for (int v = 0; v < 5; ++v) Flo[v] = F[v];
        } else if (task == 1) {
#pragma unroll
          for (int v = 0; v < 5; ++v) Fhi[v] = F[v];
        } else if (task == 2) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFj[v][tj * TK + tk] = F[v];
        } else if (task == 3) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFk[v][tj * (TK + 1) + tk] = F[v];
        } else if (tid < TK) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFj[v][jl * TK + tid] = F[v];
        } else {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFk[v][(tid - TK) * (TK + 1) + kl] = F[v];
        }
      }
    }
    __syncthreads();
    if (act) {
      const long long c = cidx(B.N, I, J, K);
      const double vol = B.vol[c];
      double r[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        double a = 0.0 - (Fhi[v] - Flo[v]);
        a = a - (sFj[v][(tj + 1) * TK + tk] - sFj[v][tj * TK + tk]);
        a = a - (sFk[v][tj * (TK + 1) + tk + 1] - sFk[v][tj * (TK + 1) + tk]);
        r[v] = a / vol;
      }
      const double* qi = qin.p[t.x];
      if (nd.n_sponge > 0) {
        // SmoothRampSponge, smooth_ramp.py:212-221
        const double xcv = B.xc[c];
        const long long mrow = (long long)i * B.n[1] + j;
        const long long mplane = (long long)B.n[0] * B.n[1];
        for (int sp = 0; sp < nd.n_sponge; ++sp) {
          const double* pr = nd.sponge[sp];
          double x_l = xcv / pr[4];
          double eta = fmin(fmax((x_l - pr[0]) / (pr[1] - pr[0]), 0.0), 1.0);
          double ramp = (eta * eta * eta) * (10.0 + eta * (-15.0 + 6.0 * eta));
          double sigma = ((pr[2] * pr[3]) / pr[4]) * ramp;
#pragma unroll
          for (int v = 0; v < 5; ++v)
            r[v] = r[v] + (-sigma) * (qi[v * cells + c] - B.sponge_mean[v * mplane + mrow]);
        }
      }
      double* o = out.p[t.x];
      if (MODE == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v * cells + c] = r[v];
      } else if (MODE == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v * cells + c] = qi[v * cells + c] + dt * r[v];
      } else {
        const double* qz = q0.p[t.x];
        const double a0 = MODE == 2 ? 0.75 : 1.0 / 3.0;
        const double a1 = MODE == 2 ? 0.25 : 2.0 / 3.0;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          o[v * cells + c] = a0 * qz[v * cells + c] + a1 * (qi[v * cells + c] + dt * r[v]);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) Flo[v] = Fhi[v];
    }
    __syncthreads();
  }
}

template __global__ void k_flux<0>(const BlockDev*, const int4*, int, NumDev, Ptrs, Ptrs, Ptrs,
                                   const double*, double);
template __global__ void k_flux<1>(const BlockDev*, const int4*, int, NumDev, Ptrs, Ptrs, Ptrs,
                                   const double*, double);
template __global__ void k_flux<2>(const BlockDev*, const int4*, int, NumDev, Ptrs, Ptrs, Ptrs,
                                   const double*, double);
template __global__ void k_flux<3>(const BlockDev*, const int4*, int, NumDev, Ptrs, Ptrs, Ptrs,
                                   const double*, double);

// ------------------------------------------------------------------- shells / sponge
// save: shell <- q ghosts
__global__ void k_shell_save(const BlockDev* __restrict__ blks, const int* __restrict__ wg_prefix,
                             int nblk, Ptrs q) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  long long s = (long long)(blockIdx.x - wg_prefix[b]) * blockDim.x + threadIdx.x;
  if (s >= B.shell_cells) return;
  int I, J, K;
  shell_cell(s, B.N, B.n, B.ng, &I, &J, &K);
  const long long c = cidx(B.N, I, J, K);
  for (int v = 0; v < 5; ++v) B.shell[v * B.shell_cells + s] = q.p[b][v * B.cells + c];
}

// q0 ghosts <- shell (restore); out ghosts <- the three stage combinations applied to
// the saved ghosts with R = 0 there (solver.py:105,110-121); zero: out ghosts <- 0.
template <int MODE>  // 0: zero out; 1: restore q0 + RK3 ghost combination into out
__global__ void k_shell_finish(const BlockDev* __restrict__ blks, const int* __restrict__ wg_prefix,
                               int nblk, Ptrs q0, Ptrs out, const double* __restrict__ d_dt,
                               double dt_host) {
  const int b = find_block(wg_prefix, nblk, blockIdx.x);
  const BlockDev& B = blks[b];
  long long s = (long long)(blockIdx.x - wg_prefix[b]) * blockDim.x + threadIdx.x;
  if (s >= B.shell_cells) return;
  int I, J, K;
  shell_cell(s, B.N, B.n, B.ng, &I, &J, &K);
  const long long c = cidx(B.N, I, J, K);
  if (MODE == 0) {
    for (int v = 0; v < 5; ++v) out.p[b][v * B.cells + c] = 0.0;
    return;
  }
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  const double zr = dt * 0.0;
  for (int v = 0; v < 5; ++v) {
    const double g = B.shell[v * B.shell_cells + s];
    q0.p[b][v * B.cells + c] = g;
    const double q1 = g + zr;
    const double q2 = 0.75 * g + 0.25 * (q1 + zr);
    out.p[b][v * B.cells + c] = (1.0 / 3.0) * g + (2.0 / 3.0) * (q2 + zr);
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

// ----------------------------------------------------------------------- launchers
hipError_t launch_wm_eval(const NumDev& nd, int mode, long long n, const double* a, const double* b,
                          const double* c, const double* d, double T_wall, double* o0, double* o1,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wm_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nd, mode, n, a,
                     b, c, d, T_wall, o0, o1);
  return hipGetLastError();
}

hipError_t launch_prim(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q, GasC gas,
                       hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prim, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, q, gas);
  return hipGetLastError();
}

hipError_t launch_grad(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, const NumDev& nd,
                       hipStream_t s) {
  if (nwg <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_grad, dim3(nwg), dim3(256), 0, s, d_blks, d_pre, nblk, nd);
  return hipGetLastError();
}

hipError_t launch_wall(const BlockDev* d_blks, const WallOp* d_ops, int nops, int maxcount, Ptrs q,
                       const NumDev& nd, hipStream_t s) {
  if (nops <= 0 || maxcount <= 0) return hipSuccess;
  dim3 grid((maxcount + 255) / 256, nops);
  hipLaunchKernelGGL(k_wall, grid, dim3(256), 0, s, d_blks, d_ops, q, nd);
  return hipGetLastError();
}

hipError_t launch_flux(int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles, int ci_len,
                       const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt,
                       double dt, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  dim3 grid(ntiles), block(TJ * TK);
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_flux<0>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    case 1: hipLaunchKernelGGL(k_flux<1>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    case 2: hipLaunchKernelGGL(k_flux<2>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    default: hipLaunchKernelGGL(k_flux<3>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
  }
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

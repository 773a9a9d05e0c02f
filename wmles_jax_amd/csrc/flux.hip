// Flux-kernel launches.  k_flux2 (FP32-bulk) runs one marching tile per workgroup of the
// body in flux2.h.  k_flux1 is the FP64 plain form: 15 x 16 (j, k) cell tiles marching i,
// every face flux evaluated once (the i face in registers, the j / k faces exchanged
// through LDS), each cell summing its six faces (reference solver.py:86-95: inviscid scheme
// + viscous flux, wall override, divergence / vol, then sources, solver.py:103-104, and
// the stage combination, solver.py:109-121).
#include "flux2.h"

namespace wmx {

// MODE 0: r = RHS; 1: q + dt r; 2: 3/4 q0 + 1/4 (q + dt r); 3: 1/3 q0 + 2/3 (q + dt r)
template <class R, int MODE, int SC>
__global__ __launch_bounds__(256, 2) void k_flux2(const BlockDev* __restrict__ blks,
                                               const int4* __restrict__ tiles, int ci_len,
                                               NumDev nd, Ptrs qin, Ptrs q0, Ptrs out,
                                               const double* __restrict__ d_dt, double dt_host) {
  flux2_tile<R, MODE, SC>(blks, tiles, ci_len, nd, qin, q0, out, d_dt, dt_host, (int)blockIdx.x);
}

// ------------------------------------------------------------------------------ flux
__device__ __forceinline__ CellW load_w(gconst<double> P, long long cells, long long c) {
  CellW w;
  w.r = P[c];
  w.u = P[cells + c];
  w.v = P[2 * cells + c];
  w.w = P[3 * cells + c];
  w.p = P[4 * cells + c];
  w.e = P[7 * cells + c];
  return w;
}

__device__ __forceinline__ CellG load_g(gconst<double> G, long long cells, long long c) {
  CellG g;
#pragma unroll
  for (int i = 0; i < 12; ++i) g.g[i] = G[i * cells + c];
  g.mu = G[G_MU * cells + c];
  g.k = G[G_K * cells + c];
  return g;
}

// FP64 tile staging of one plane (k_flux1): gradient records on the tile + 1-cell ring
// (17 x 18 cells, unpadded) and primitive records (r u v w p e T) on the tile + 2-cell ring
// (19 x 20 cells)
constexpr int NPW64 = 7;
struct Stage64 {
  const double (*sG)[NG2] = nullptr;   // [NGR][NG2]
  int gL = -1, gR = -1;                // gradient slots of the left / right cell (-1: workspace)
  const double (*sP)[NP2] = nullptr;   // [NPW64][NP2]
  int pR = -1, pst = 0;                // primitive slot of the right cell (-1: workspace), stride
};

__device__ __forceinline__ CellW sp_w64(const double (*sP)[NP2], int s) {
  CellW w;
  w.r = sP[0][s]; w.u = sP[1][s]; w.v = sP[2][s]; w.w = sP[3][s]; w.p = sP[4][s]; w.e = sP[5][s];
  return w;
}

__device__ __forceinline__ CellG sg_g64(const double (*sG)[NG2], int s) {
  CellG g;
#pragma unroll
  for (int i = 0; i < 12; ++i) g.g[i] = sG[i][s];
  g.mu = sG[G_MU][s];
  g.k = sG[G_K][s];
  return g;
}

// Numerical flux (times area) through face f of direction d, f = coordinate of the
// right cell (I, J, K); solver.py:88-93 with the wall override.  st: the tile's staged
// records of the face's cells where it has them (else the workspace is read).
__device__ __forceinline__ void face_flux(const NumDev& nd, const BlockDev& B, int d, int I, int J,
                                          int K, double* F, const Stage64& st = Stage64()) {
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
    const gconst<double> W = gmem(B.wall[2 * d + side]);
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = W[v * nw + w];
    return;
  }
  const int* N = B.N;
  const long long cells = B.cells;
  const long long sd = d == 0 ? (long long)N[1] * N[2] : (d == 1 ? (long long)N[2] : 1LL);
  const long long cR = cidx(N, I, J, K), cL = cR - sd;
  const gconst<double> P = gmem(B.P);
  const gconst<double> G = gmem(B.G);
  const long long fi = fidx(N, d, I, J, K), fc = fcount(N, d);
  const gconst<double> S = gmem(B.S[d]);
  const double Sx = S[fi], Sy = S[fc + fi], Sz = S[2 * fc + fi];
  const bool sp = st.pR >= 0;
  const CellW L = sp ? sp_w64(st.sP, st.pR - st.pst) : load_w(P, cells, cL);
  const CellW R = sp ? sp_w64(st.sP, st.pR) : load_w(P, cells, cR);
  const int sch = nd.scheme;
  if (sch == WMX_SCHEME_KEP || sch == WMX_SCHEME_KEEP) {
    CellW L2 = L, R2 = R;
    if (nd.order == 4) {
      L2 = sp ? sp_w64(st.sP, st.pR - 2 * st.pst) : load_w(P, cells, cL - sd);
      R2 = sp ? sp_w64(st.sP, st.pR + st.pst) : load_w(P, cells, cR + sd);
    }
    if (sch == WMX_SCHEME_KEP) central_flux<false>(nd.order, L2, L, R, R2, Sx, Sy, Sz, F);
    else central_flux<true>(nd.order, L2, L, R, R2, Sx, Sy, Sz, F);
  } else {
    const CellW L2 = sp ? sp_w64(st.sP, st.pR - 2 * st.pst) : load_w(P, cells, cL - sd);
    const CellW R2 = sp ? sp_w64(st.sP, st.pR + st.pst) : load_w(P, cells, cR + sd);
    if (sch == WMX_SCHEME_MUSCLHLLC) {
      muscl_hllc_flux(nd.gas.gamma, nd.limiter, L2, L, R, R2, Sx, Sy, Sz, F);
    } else {
      // Hybrid: (1 - s) F_c + s F_u; a part whose weight is exactly 0 is not
      // evaluated (bit-identical while the skipped flux is finite).
      const double thL = st.gL >= 0 ? st.sG[G_TH][st.gL] : G[G_TH * cells + cL];
      const double thR = st.gR >= 0 ? st.sG[G_TH][st.gR] : G[G_TH * cells + cR];
      const double th = fmax(thL, thR);
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
  const CellG gL = st.gL >= 0 ? sg_g64(st.sG, st.gL) : load_g(G, cells, cL);
  const CellG gR = st.gR >= 0 ? sg_g64(st.sG, st.gR) : load_g(G, cells, cR);
  const double phL[4] = {L.u, L.v, L.w, sp ? st.sP[6][st.pR - st.pst] : P[P_T * cells + cL]};
  const double phR[4] = {R.u, R.v, R.w, sp ? st.sP[6][st.pR] : P[P_T * cells + cR]};
  const gconst<double> XC = gmem(B.xc);
  const double xL[3] = {XC[cL], XC[cells + cL], XC[2 * cells + cL]};
  const double xR[3] = {XC[cR], XC[cells + cR], XC[2 * cells + cR]};
  viscous_face(gL, gR, phL, phR, xL, xR, Sx, Sy, Sz, F);
}

// FP64 plain form: every face reads its stencil through the caches (no register
// window), which keeps the kernel near 200 VGPRs / 2 waves per SIMD.  Tiles of 15 x 16
// cells on a 16 x 16 thread grid, as in k_flux2: per plane thread (tj, tk) evaluates the
// j face below row tj, the k face left of its cell (tj < 15) or the upper edge k face of
// row tk (tj = 15), and the i face above its cell (tj < 15) -- three face tasks for every
// wavefront.
// MODE 0: r = RHS; 1: q + dt r; 2: 3/4 q0 + 1/4 (q + dt r); 3: 1/3 q0 + 2/3 (q + dt r)
template <int MODE>
__global__ __launch_bounds__(256) void k_flux1(const BlockDev* __restrict__ blks,
                                              const int4* __restrict__ tiles, int ci_len,
                                              NumDev nd, Ptrs qin, Ptrs q0, Ptrs out,
                                              const double* __restrict__ d_dt, double dt_host) {
  __shared__ double sFj[5][(C2J + 1) * TK];
  __shared__ double sFk[5][C2J * (TK + 1)];
  __shared__ double sG[NGR][NG2];                // gradient records of plane I, tile + 1-cell ring
  __shared__ double sP[NPW64][NP2];              // primitives of plane I, tile + 2-cell ring
  const int4 t = tiles[blockIdx.x];
  const BlockDev& B = blks[t.x];
  const int j0 = t.y, k0 = t.z, i0 = t.w;
  const int i1 = min(i0 + ci_len, B.n[0]);
  const int tid = threadIdx.x;
  const int tj = tid / TK, tk = tid % TK;
  const int j = j0 + tj, k = k0 + tk;
  const bool own = tj < C2J;
  const bool act = own && j < B.n[1] && k < B.n[2];
  const int jl = min(j0 + C2J, B.n[1]) - j0;
  const int kl = min(k0 + TK, B.n[2]) - k0;
  const int ng = B.ng;
  const int J = ng + j, K = ng + k;
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  const long long cells = B.cells;
  double Flo[5], Fhi[5];
  if (act) face_flux(nd, B, 0, ng + i0, J, K, Flo);
  const gconst<double> Gw = gmem(B.G);
  const gconst<double> Pw = gmem(B.P);
  const int pl[NPW64] = {P_R, P_U, P_V, P_W, P_P, P_E, P_T};
  for (int i = i0; i < i1; ++i) {
    const int I = ng + i;
    // wavefront priorities by phase: the staging (loads whose round trip the barrier below
    // waits for) 2, the face tasks 1, the epilogue 0; of the two workgroups on a CU, the
    // one nearer its next barrier issues first
    __builtin_amdgcn_s_setprio(2);
    // plane I staged for the j / k faces and the lower cell of the i face: the gradient
    // records on the tile + 1-cell ring (rows j0-1 .. j0+15, columns k0-1 .. k0+16, within
    // the cells the gradient launch covers) and the primitives on the tile + 2-cell ring
    // (within the array)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int r = tid + u * NT;
      if (r < NG2) {
        const int gj = r / G2K, gk = r - gj * G2K;
        const int JJ = ng + j0 - 1 + gj, KK = ng + k0 - 1 + gk;
        if (JJ <= ng + B.n[1] && KK <= ng + B.n[2]) {
          const long long c = cidx(B.N, I, JJ, KK);
          double v[NGR];
#pragma unroll
          for (int x = 0; x < NGR; ++x) v[x] = Gw[x * cells + c];
#pragma unroll
          for (int x = 0; x < NGR; ++x) sG[x][r] = v[x];
        }
      }
      const int rp = tid + u * NT;
      if (rp < NP2) {
        const int pj = rp / H2K, pk = rp - pj * H2K;
        const int JJ = ng + j0 - 2 + pj, KK = ng + k0 - 2 + pk;
        if (JJ < B.N[1] && KK < B.N[2] && JJ >= 0 && KK >= 0) {
          const long long c = cidx(B.N, I, JJ, KK);
          double v[NPW64];
#pragma unroll
          for (int x = 0; x < NPW64; ++x) v[x] = Pw[pl[x] * cells + c];
#pragma unroll
          for (int x = 0; x < NPW64; ++x) sP[x][rp] = v[x];
        }
      }
    }
    __builtin_amdgcn_s_setprio(1);
    __syncthreads();
#pragma unroll 1
    for (int task = 0; task < 3; ++task) {
      int d, fI = I, rj = tj, rk = tk;
      bool doit;
      if (task == 0) { d = 1; doit = rj <= jl && rk < kl; }
      else if (task == 1) {
        d = 2;
        if (own) doit = rj < jl && rk < kl;
        else { rj = tk; rk = kl; doit = rj < jl; }
      } else { d = 0; fI = I + 1; doit = act; }
      double F[5];
      if (doit) {
        // staged slots of the face's cells (the i face's stencil reaches planes I - 1 .. I + 2:
        // its primitives and its upper cell's gradient record come from the workspace)
        Stage64 st;
        st.sG = sG;
        st.sP = sP;
        if (d == 0) {
          st.gL = (tj + 1) * G2K + (tk + 1);
        } else {
          st.gR = (rj + 1) * G2K + (rk + 1);
          st.gL = d == 1 ? rj * G2K + (rk + 1) : (rj + 1) * G2K + rk;
          st.pR = (rj + 2) * H2K + (rk + 2);
          st.pst = d == 1 ? H2K : 1;
        }
        face_flux(nd, B, d, fI, ng + j0 + rj, ng + k0 + rk, F, st);
        if (task == 0) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFj[v][rj * TK + rk] = F[v];
        } else if (task == 1) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sFk[v][rj * (TK + 1) + rk] = F[v];
        } else {
#pragma unroll
          for (int v = 0; v < 5; ++v) Fhi[v] = F[v];
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
    // the epilogue's state and volume loads go out before the barrier
    const long long c = cidx(B.N, I, J, K);
    const gconst<double> qi = gmem(qin.p[t.x]);
    double vq[5], vz[5], vol = 1.0;
    if (act) {
      vol = gmem(B.vol)[c];
#pragma unroll
      for (int v = 0; v < 5; ++v) vq[v] = qi[v * cells + c];
      if (MODE >= 2) {
        const gconst<double> qz = gmem(q0.p[t.x]);
#pragma unroll
        for (int v = 0; v < 5; ++v) vz[v] = qz[v * cells + c];
      }
    }
    __syncthreads();
    if (act) {
      double r[5];
      const double ivol = 1.0 / vol;             // (within an ulp of dividing each component)
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        double a = 0.0 - (Fhi[v] - Flo[v]);
        a = a - (sFj[v][(tj + 1) * TK + tk] - sFj[v][tj * TK + tk]);
        a = a - (sFk[v][tj * (TK + 1) + tk + 1] - sFk[v][tj * (TK + 1) + tk]);
        r[v] = a * ivol;
      }
      if (nd.n_sponge > 0) {
        // SmoothRampSponge, smooth_ramp.py:212-221
        const double xcv = gmem(B.xc)[c];
        const long long mrow = (long long)i * B.n[1] + j;
        const long long mplane = (long long)B.n[0] * B.n[1];
        for (int sp = 0; sp < nd.n_sponge; ++sp) {
          // a zero weight adds -0 * (q - mean): skipped (rows outside every sponge have no mean)
          const double sigma = sponge_sigma(nd.sponge[sp], xcv);
          if (sigma != 0.0) {
#pragma unroll
            for (int v = 0; v < 5; ++v)
              r[v] = r[v] + (-sigma) * (vq[v] - gmem(B.sponge_mean)[v * mplane + mrow]);
          }
        }
      }
      const gmut<double> o = gmem_w(out.p[t.x]);
      if (MODE == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v * cells + c] = r[v];
      } else if (MODE == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v * cells + c] = vq[v] + dt * r[v];
      } else {
        const double a0 = MODE == 2 ? 0.75 : 1.0 / 3.0;
        const double a1 = MODE == 2 ? 0.25 : 2.0 / 3.0;
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v * cells + c] = a0 * vz[v] + a1 * (vq[v] + dt * r[v]);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) Flo[v] = Fhi[v];
    }
    __syncthreads();
  }
}

// The FP64 kernels and the Hybrid (KEP4 + MUSCL-HLLC) FP32-bulk kernels are compiled in
// flux_ilp.hip (this file with WMX_FLUX_ILP_UNIT defined) under LLVM's max-ILP machine
// scheduler, which shortens them by 2-3% (the arithmetic is the same); the other FP32-bulk
// kernels, which it slows, are instantiated here by their launches.
#define WMX_FLUX_SIG (const BlockDev*, const int4*, int, NumDev, Ptrs, Ptrs, Ptrs, const double*, double)
#define WMX_FLUX_ILP_LIST(PFX)                                                   \
  PFX template __global__ void k_flux1<0> WMX_FLUX_SIG;                          \
  PFX template __global__ void k_flux1<1> WMX_FLUX_SIG;                          \
  PFX template __global__ void k_flux1<2> WMX_FLUX_SIG;                          \
  PFX template __global__ void k_flux1<3> WMX_FLUX_SIG;                          \
  PFX template __global__ void k_flux2<float, 0, SC_HYB_KEP4_ALBADA> WMX_FLUX_SIG; \
  PFX template __global__ void k_flux2<float, 1, SC_HYB_KEP4_ALBADA> WMX_FLUX_SIG; \
  PFX template __global__ void k_flux2<float, 2, SC_HYB_KEP4_ALBADA> WMX_FLUX_SIG; \
  PFX template __global__ void k_flux2<float, 3, SC_HYB_KEP4_ALBADA> WMX_FLUX_SIG;

#ifdef WMX_FLUX_ILP_UNIT
WMX_FLUX_ILP_LIST()
#else
WMX_FLUX_ILP_LIST(extern)


hipError_t launch_flux(int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles, int ci_len,
                       const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt,
                       double dt, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  dim3 grid(ntiles), block(NT);
  if (nd.prec == WMX_PREC_FP32_BULK) {
#define WMX_F2(M, C) hipLaunchKernelGGL((k_flux2<float, M, C>), grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt)
#define WMX_F2M(C)              \
  switch (mode) {               \
    case 0: WMX_F2(0, C); break; \
    case 1: WMX_F2(1, C); break; \
    case 2: WMX_F2(2, C); break; \
    default: WMX_F2(3, C); break; \
  }
    switch (scheme_config(nd)) {
      case SC_HYB_KEP4_ALBADA: WMX_F2M(SC_HYB_KEP4_ALBADA); break;
      case SC_KEP2: WMX_F2M(SC_KEP2); break;
      case SC_KEP4: WMX_F2M(SC_KEP4); break;
      default: WMX_F2M(SC_ANY); break;
    }
#undef WMX_F2M
#undef WMX_F2
    return hipGetLastError();
  }
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_flux1<0>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    case 1: hipLaunchKernelGGL(k_flux1<1>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    case 2: hipLaunchKernelGGL(k_flux1<2>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
    default: hipLaunchKernelGGL(k_flux1<3>, grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt); break;
  }
  return hipGetLastError();
}
#endif

}  // namespace wmx

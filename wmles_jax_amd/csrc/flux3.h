// FP32-bulk fused step kernel body with the gradients computed in the march (flux3_tile):
// Green-Gauss gradients, SGS viscosity and Ducros sensor of every plane are evaluated in
// LDS from the staged primitive records, so the gradient records never go through HBM and
// the separate gradient launch is not needed (reference viscous.py:6-25, sgs.py,
// solver.py:78-95 for the composition; the face fluxes, divergence and stage update are
// those of flux2_tile).
//
// Skewed march.  At plane I of a tile's i chunk the workgroup
//   1. stages (u, v, w, T') of plane I+1 (three rolling slots hold planes I-1, I, I+1),
//      (rho, p', mu, e') and the centroids of plane I, own columns from registers loaded
//      one plane ahead, the halo ring from loads issued during the previous plane;
//   2. evaluates the gradient records of plane I on the tile + 1-cell ring (306 cells; at
//      the chunk's two end planes, which feed only the i faces, the 240 own cells);
//   3. evaluates the j / k faces of plane I (exchanged through LDS) and the i face I-1/2 of
//      the own column (registers);
//   4. finishes cell I-1: its i faces I-3/2 and I-1/2 (registers) and its j / k faces,
//      kept from the previous plane in the other half of the double-buffered exchange.
// The march runs over planes I0-1 .. I0+len, one plane more at each end than the tile's
// cells.  Three barriers per plane (staging, gradients, face exchange; the last is skipped
// at the end planes, which have no j / k faces).
//
// Wavefront priorities (s_setprio) by phase: two workgroups share each CU, one wavefront of
// each per SIMD, and the launch is bound by latency at the barriers, not by issue.  A
// wavefront in its gradient phase issues first (its workgroup's next barrier waits for the
// slowest wavefront's gradient cells, whose geometry loads carry a memory round trip), then
// one in its face tasks; the epilogue and the staging run at the base priority.
#pragma once
#include "flux2.h"

namespace wmx {

namespace {

// wavefront priorities of the phases of a plane (s_setprio operands, 0..3)
constexpr int PRIO_GRAD = 2, PRIO_FACE = 1, PRIO_REST = 0;

// a cell's FP32-bulk primitive record incl. mu: groups (u, v, w, T') and (rho, p', mu, e')
struct Rec32 {
  float u, v, w, T, r, p, mu, e;
};

__device__ __forceinline__ Rec32 load_rec32(const float* P, size_t cells, size_t c) {
  const float4 a = ld4(P + 4 * c), b = ld4(P + 4 * (cells + c));
  Rec32 o;
  o.u = a.x; o.v = a.y; o.w = a.z; o.T = a.w;
  o.r = b.x; o.p = b.y; o.mu = b.z; o.e = b.w;
  return o;
}

__device__ __forceinline__ CellWT<float> rec_w(const Rec32& a) {
  CellWT<float> w;
  w.r = a.r; w.u = a.u; w.v = a.v; w.w = a.w; w.p = a.p; w.e = a.e;
  return w;
}

// Gradient record of cell (I, J, K) (P-region slot ps, gradient-region slot gs) from the
// staged planes: the per-cell arithmetic of k_grad32 (grad32.h), faces in the order i, j, k.
__device__ __forceinline__ void grad_cell3(const BlockDev& B, const NumDev& nd, const float (*pm)[NP2],
                                           const float (*p0)[NP2], const float (*pp)[NP2],
                                           const float (*pr)[NP2], int ps, int I, int J, int K, float* G,
                                           float* Slo) {
  const int N1 = B.N[1], N2 = B.N[2];
  const unsigned st0 = (unsigned)N1 * N2;
  const unsigned fc0 = (unsigned)fcount(B.N, 0), fc1 = (unsigned)fcount(B.N, 1), fc2 = (unsigned)fcount(B.N, 2);
  const unsigned c = (unsigned)I * st0 + (unsigned)J * N2 + (unsigned)K;
  const unsigned f0 = c;
  const unsigned f1 = (unsigned)I * ((unsigned)(N1 + 1) * N2) + (unsigned)J * N2 + (unsigned)K;
  const unsigned f2 = (unsigned)I * ((unsigned)N1 * (N2 + 1)) + (unsigned)J * (unsigned)(N2 + 1) + (unsigned)K;
  const float* S0 = B.Sf[0];
  const float* S1 = B.Sf[1];
  const float* S2 = B.Sf[2];
  // every load first (one memory round trip), then the arithmetic
  float S0lo[3], S0hi[3], S1lo[3], S1hi[3], S2lo[3], S2hi[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    S0lo[x] = ldu(S0 + x * (size_t)fc0, f0);
    S0hi[x] = ldu(S0 + x * (size_t)fc0, f0 + st0);
    S1lo[x] = ldu(S1 + x * (size_t)fc1, f1);
    S1hi[x] = ldu(S1 + x * (size_t)fc1, f1 + (unsigned)N2);
    S2lo[x] = ldu(S2 + x * (size_t)fc2, f2);
    S2hi[x] = ldu(S2 + x * (size_t)fc2, f2 + 1);
  }
  const float volr = ldu((const float*)B.volf, c);
  const float h = 0.5f;
  float ph0[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) ph0[v] = p0[v][ps];
  float acc[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float flo_v = h * (pm[v][ps] + ph0[v]);
    const float fhi_v = h * (ph0[v] + pp[v][ps]);
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[3 * v + x] = fhi_v * S0hi[x] - flo_v * S0lo[x];
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float flo_v = h * (p0[v][ps - H2K] + ph0[v]);
    const float fhi_v = h * (ph0[v] + p0[v][ps + H2K]);
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * S1hi[x] - flo_v * S1lo[x]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float flo_v = h * (p0[v][ps - 1] + ph0[v]);
    const float fhi_v = h * (ph0[v] + p0[v][ps + 1]);
#pragma unroll
    for (int x = 0; x < 3; ++x) acc[3 * v + x] = acc[3 * v + x] + (fhi_v * S2hi[x] - flo_v * S2lo[x]);
  }
  float g[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) g[i] = fdiv<float>(acc[i], volr);
  const float rho = pr[0][ps], mu = pr[2][ps];
  float mut = 0.0f * rho;
  if (nd.sgs != WMX_SGS_NONE) {
    // filter width vol^(1/3) (solver.py:82), single-precision cube root
    const float delta = cbrtf(volr);
    const float cs = (float)nd.sgs_c;
    float nut;
    if (nd.sgs == WMX_SGS_SMAGORINSKY) nut = nut_smagorinsky<float>(g, delta, cs);
    else if (nd.sgs == WMX_SGS_WALE) nut = nut_wale<float>(g, delta, cs);
    else nut = nut_vreman<float>(g, delta, cs);
    mut = rho * nut;
  }
  const GasC& gs = nd.gas;
#pragma unroll
  for (int i = 0; i < 12; ++i) G[i] = g[i];
  G[12] = mu + mut;
  G[13] = (float)gs.cp * (fdiv<float>(mu, (float)gs.Pr) + fdiv<float>(mut, (float)gs.Pr_t));
  G[14] = ducros<float>(g);
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    Slo[x] = S0lo[x];
    Slo[3 + x] = S1lo[x];
    Slo[6 + x] = S2lo[x];
  }
}

}  // namespace

// One marching tile of the fused FP32-bulk launch; called by every thread of a 256-thread
// workgroup.  MODE 0: r = RHS; 1: q + dt r; 2: 3/4 q0 + 1/4 (q + dt r); 3: 1/3 q0 + 2/3
// (q + dt r)
template <int MODE, int SC>
__device__ __forceinline__ void flux3_tile(const BlockDev* __restrict__ blks, const int4* __restrict__ tiles,
                                           int ci_len, const NumDev& nd, const Ptrs& qin, const Ptrs& q0,
                                           const Ptrs& out, const double* __restrict__ d_dt, double dt_host,
                                           int tile) {
  __shared__ float sPh[3][4][NP2];               // (u, v, w, T'), planes in rolling slots
  __shared__ float sPr[4][NP2];                  // (rho, p', mu, e') of the current plane
  __shared__ float sG[NGW][NG2S];                // gradient records of the current plane
  __shared__ double sX[3][NG2S];                 // centroids of the current plane
  __shared__ float sS[3][3][NG2S];               // lower i / j / k face vectors of the gradient cells
  __shared__ float sFj[NFX][NT];                 // slot tile_tid(r, c)
  __shared__ float sFk[NFX][NT + C2J + 1];       // slot kface_slot(r, c)
  const int4 t = tiles[tile];
  const BlockDev& B = blks[t.x];
  const int j0 = t.y, k0 = t.z, i0 = t.w;
  const int i1 = min(i0 + ci_len, B.n[0]);
  const int tid = threadIdx.x;
  const int tj = (tid >> 6) + 4 * ((tid >> 4) & 3), tk = tid & 15;   // tile_tid(tj, tk) == tid
  const int ng = B.ng;
  const int* N = B.N;
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int j = j0 + tj, k = k0 + tk;
  const bool own = tj < C2J;
  const bool act = own && j < n1 && k < n2;
  const int jl = min(j0 + C2J, n1) - j0;
  const int kl = min(k0 + TK, n2) - k0;
  const int J = ng + j, K = ng + k;
  const unsigned cells = (unsigned)B.cells;
  const unsigned sI = (unsigned)N[1] * N[2];
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  const float* P = (const float*)B.Pf;
  const double* XC = B.xc;
  auto p_ok = [&](int jj, int kk) { return jj >= 0 && jj < N[1] && kk >= 0 && kk < N[2]; };
  auto g_ok = [&](int jj, int kk) {
    return jj >= ng - 1 && jj <= ng + n1 && kk >= ng - 1 && kk <= ng + n2;
  };
  const int I0 = ng + i0, Iend = ng + i1;
  const int base = I0 - 2;                       // plane p lives in slot (p - base) % 3
  auto slot = [&](int p) { return (p - base) % 3; };
  // own column (clamped into the array for threads off it) and its slots
  const unsigned cb = (unsigned)min(J, N[1] - 1) * N[2] + min(K, N[2] - 1);
  const bool own_p = own && p_ok(J, K), own_g = own && g_ok(J, K);
  const int sp_own = (tj + 2) * H2K + (tk + 2), sg_own = (tj + 1) * G2P + (tk + 1);
  // halo ring job (plane-independent): 140 primitive slots (halo 2) or 66 centroid slots
  // (halo 1)
  constexpr int NRP = NP2 - C2J * TK, NRG = NG2 - C2J * TK;
  int ring = 0, ring_s = 0;
  unsigned ring_off = 0;
  if (tid < NRP) {
    int sj, sk;
    ring_slot(tid, 2, C2J, &sj, &sk);
    const int jj = ng + j0 - 2 + sj, kk = ng + k0 - 2 + sk;
    if (p_ok(jj, kk)) { ring = 1; ring_s = sj * H2K + sk; ring_off = (unsigned)jj * N[2] + kk; }
  } else if (tid < NRP + NRG) {
    int sj, sk;
    ring_slot(tid - NRP, 1, C2J, &sj, &sk);
    const int jj = ng + j0 - 1 + sj, kk = ng + k0 - 1 + sk;
    if (g_ok(jj, kk)) { ring = 2; ring_s = sj * G2P + sk; ring_off = (unsigned)jj * N[2] + kk; }
  }
  // ring registers: primitive slots (u v w T') of I+1 in rva and (r p' mu e') of I in rvb;
  // centroid slots the centroid of I, its three doubles in the bits of rva.xy, rva.zw, rvb.xy
  float4 rva = make_float4(0.f, 0.f, 0.f, 0.f), rvb = rva;
  auto ring_load = [&](int Iphi, int Irec) {                  // for the plane staged next
    if (ring == 1) {
      rva = ld4(P + 4 * ((size_t)Iphi * sI + ring_off));
      rvb = ld4(P + 4 * ((size_t)cells + (size_t)Irec * sI + ring_off));
    } else if (ring == 2) {
      const unsigned o = (unsigned)Irec * sI + ring_off;
      const double x0 = ldu(XC, o), x1 = ldu(XC + (size_t)cells, o), x2 = ldu(XC + 2 * (size_t)cells, o);
      rva.x = __int_as_float(__double2loint(x0)); rva.y = __int_as_float(__double2hiint(x0));
      rva.z = __int_as_float(__double2loint(x1)); rva.w = __int_as_float(__double2hiint(x1));
      rvb.x = __int_as_float(__double2loint(x2)); rvb.y = __int_as_float(__double2hiint(x2));
    }
  };
  auto ring_x = [&](int x) {
    const float lo = x == 0 ? rva.x : (x == 1 ? rva.z : rvb.x);
    const float hi = x == 0 ? rva.y : (x == 1 ? rva.w : rvb.y);
    return __hiloint2double(__float_as_int(hi), __float_as_int(lo));
  };
  // ---- prologue: planes I0-2 and I0-1 of (u v w T') staged; windows and prefetches for I0-1
  Rec32 Pw[4], Pn;
  Pw[2] = load_rec32(P, cells, (size_t)(I0 - 2) * sI + cb);
  Pw[3] = load_rec32(P, cells, (size_t)(I0 - 1) * sI + cb);
  Pn = load_rec32(P, cells, (size_t)I0 * sI + cb);
  double Xp[3], Xn[3];                           // own centroids of I-1 and (prefetched) I
#pragma unroll
  for (int x = 0; x < 3; ++x) Xn[x] = ldu(XC + x * (size_t)cells, (unsigned)(I0 - 1) * sI + cb);
  if (ring == 1) {
    const float4 a2 = ld4(P + 4 * ((size_t)(I0 - 2) * sI + ring_off));
    const float4 a1 = ld4(P + 4 * ((size_t)(I0 - 1) * sI + ring_off));
    const float v2[4] = {a2.x, a2.y, a2.z, a2.w}, v1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sPh[slot(I0 - 2)][v][ring_s] = v2[v];
      sPh[slot(I0 - 1)][v][ring_s] = v1[v];
    }
  }
  if (own_p) {
    const float v2[4] = {Pw[2].u, Pw[2].v, Pw[2].w, Pw[2].T}, v1[4] = {Pw[3].u, Pw[3].v, Pw[3].w, Pw[3].T};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sPh[slot(I0 - 2)][v][sp_own] = v2[v];
      sPh[slot(I0 - 1)][v][sp_own] = v1[v];
    }
  }
  ring_load(I0, I0 - 1);
  float Gp[NGW];                                 // gradient record of the own cell at I-1
  // partial outward sum of the own cell I-1 in FP64: its lower i face minus its j / k face
  // sums (from the plane's exchange); the upper i face is subtracted when it is evaluated
  double Cp[NFX];
#pragma unroll
  for (int v = 0; v < NFX; ++v) Cp[v] = 0.0;
#pragma unroll
  for (int x = 0; x < 3; ++x) Xp[x] = 0.0;
#pragma unroll
  for (int v = 0; v < NGW; ++v) Gp[v] = 0.f;
  for (int I = I0 - 1; I <= Iend; ++I) {
    const bool inner = I >= I0 && I < Iend;      // plane owned by the tile: j / k faces
    // ---- window shift and staging of (u v w T') at I+1, (r p' mu e') and centroids at I
    Pw[0] = Pw[1]; Pw[1] = Pw[2]; Pw[2] = Pw[3]; Pw[3] = Pn;
    const int s_m = slot(I - 1), s_0 = slot(I), s_p = slot(I + 1);
    if (own_p) {
      sPh[s_p][0][sp_own] = Pw[3].u; sPh[s_p][1][sp_own] = Pw[3].v;
      sPh[s_p][2][sp_own] = Pw[3].w; sPh[s_p][3][sp_own] = Pw[3].T;
      sPr[0][sp_own] = Pw[2].r; sPr[1][sp_own] = Pw[2].p; sPr[2][sp_own] = Pw[2].mu; sPr[3][sp_own] = Pw[2].e;
    }
    if (own_g) {
#pragma unroll
      for (int x = 0; x < 3; ++x) sX[x][sg_own] = Xn[x];
    }
    if (ring == 1) {
      sPh[s_p][0][ring_s] = rva.x; sPh[s_p][1][ring_s] = rva.y; sPh[s_p][2][ring_s] = rva.z; sPh[s_p][3][ring_s] = rva.w;
      sPr[0][ring_s] = rvb.x; sPr[1][ring_s] = rvb.y; sPr[2][ring_s] = rvb.z; sPr[3][ring_s] = rvb.w;
    } else if (ring == 2) {
#pragma unroll
      for (int x = 0; x < 3; ++x) sX[x][ring_s] = ring_x(x);
    }
    __syncthreads();
    __builtin_amdgcn_s_setprio(PRIO_GRAD);
    // ---- gradient records of plane I: the tile + ring (306 cells) at owned planes, the own
    //      cells at the two end planes
    {
      int rr[2] = {-1, -1};
      if (inner) {
        rr[0] = tid;
        rr[1] = tid + NT < NG2 ? tid + NT : -1;
      } else if (own) {
        rr[0] = (tj + 1) * G2K + (tk + 1);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int r = rr[u];
        if (r < 0) continue;
        const int gj = r / G2K, gk = r - gj * G2K;
        const int JJ = ng + j0 - 1 + gj, KK = ng + k0 - 1 + gk;
        if (!g_ok(JJ, KK)) continue;
        float Gv[NGW], Slo[9];
        grad_cell3(B, nd, sPh[s_m], sPh[s_0], sPh[s_p], sPr, (gj + 1) * H2K + (gk + 1), I, JJ, KK, Gv, Slo);
        const int gsl = gj * G2P + gk;
#pragma unroll
        for (int v = 0; v < NGW; ++v) sG[v][gsl] = Gv[v];
#pragma unroll
        for (int v = 0; v < 9; ++v) sS[v / 3][v % 3][gsl] = Slo[v];
      }
    }
    __builtin_amdgcn_s_setprio(PRIO_FACE);
    __syncthreads();
    // ---- face fluxes: 0 j face below row tj, 1 k face left of column tk (or the upper
    //      edge k face of row tk), 2 the own i face I-1/2
    // ---- cell I-1: its state and volume loads go out before the face tasks, whose
    //      arithmetic covers their latency
    const bool fin = act && I >= I0 + 1;
    const unsigned c = (unsigned)(I - 1) * sI + (unsigned)J * N[2] + K;
    const double* qi = qin.p[t.x];
    double vq[5], vz[5], ivol = 0.0;
    if (fin) {
      ivol = ldu(B.vol, c);
#pragma unroll
      for (int v = 0; v < 5; ++v) vq[v] = ldu(qi + v * cells, c);
    }
    float Fin[NFX] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int task = inner ? 0 : 2; task < (I >= I0 ? 3 : 0); ++task) {
      bool doit = act;
      int d = 0, fI = I, fJ = J, fK = K;
      int rj = tj, rk = tk;
      CellWT<float> L2, L, Rc, R2;
      CellGT<float> gL, gR;
      float thL, thR, TL, TR;
      float dx[3] = {0.f, 0.f, 0.f};           // centroid difference, FP64 then rounded
      if (task == 2) {
        L2 = rec_w(Pw[0]); L = rec_w(Pw[1]); Rc = rec_w(Pw[2]); R2 = rec_w(Pw[3]);
        TL = Pw[1].T; TR = Pw[2].T;
#pragma unroll
        for (int i = 0; i < 12; ++i) gL.g[i] = Gp[i];
        gL.mu = Gp[12]; gL.k = Gp[13]; thL = Gp[14];
        gR = sg_g<float>(sG, sg_own); thR = sG[14][sg_own];
#pragma unroll
        for (int x = 0; x < 3; ++x) dx[x] = float(sX[x][sg_own] - Xp[x]);
      } else {
        if (task == 0) {
          d = 1;
          doit = rj <= jl && rk < kl;
        } else if (own) {
          d = 2;
          doit = rj < jl && rk < kl;
        } else {
          d = 2;
          rj = tk; rk = kl;
          doit = rj < jl;
        }
        fJ = ng + j0 + rj;
        fK = ng + k0 + rk;
        const int dj = d == 1, dk = d == 2;
        const int pR = (rj + 2) * H2K + (rk + 2), gR_ = (rj + 1) * G2P + (rk + 1);
        const int pst = dj * H2K + dk, gst = dj * G2P + dk;
        if (doit) {
          auto cw = [&](int s) {
            CellWT<float> w;
            w.r = sPr[0][s]; w.u = sPh[s_0][0][s]; w.v = sPh[s_0][1][s]; w.w = sPh[s_0][2][s];
            w.p = sPr[1][s]; w.e = sPr[3][s];
            return w;
          };
          L2 = cw(pR - 2 * pst); L = cw(pR - pst); Rc = cw(pR); R2 = cw(pR + pst);
          TL = sPh[s_0][3][pR - pst]; TR = sPh[s_0][3][pR];
          gL = sg_g<float>(sG, gR_ - gst); gR = sg_g<float>(sG, gR_);
          thL = sG[14][gR_ - gst]; thR = sG[14][gR_];
#pragma unroll
          for (int x = 0; x < 3; ++x) dx[x] = float(sX[x][gR_] - sX[x][gR_ - gst]);
        }
      }
      float F[NFX] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (doit) {
        const int cf = d == 0 ? fI : (d == 1 ? fJ : fK);
        int side = -1;
        if (cf == ng) side = 0;
        else if (cf == ng + B.n[d]) side = 1;
        if (side >= 0 && B.faces[2 * d + side].kind == WMX_BC_WALLMODEL) {
          int t1, t2;
          tangents(d, &t1, &t2);
          const int cc[3] = {fI, fJ, fK};
          const unsigned nw = (unsigned)B.n[t1] * B.n[t2];
          const unsigned w = (unsigned)(cc[t1] - ng) * B.n[t2] + (cc[t2] - ng);
          const double* W = B.wall[2 * d + side];   // gauge form (p - p_ref) S, no mass flux
#pragma unroll
          for (int v = 0; v < 5; ++v) F[v] = (float)ldu(W + v * nw, w);
        } else {
          // the face vector, staged by the gradient cell whose lower face it is
          const int ss = task == 2 ? sg_own : (rj + 1) * G2P + (rk + 1);
          flux_core_f32<SC>(nd, L2, L, Rc, R2, thL, thR, gL, gR, TL, TR, dx[0], dx[1], dx[2], sS[d][0][ss], sS[d][1][ss],
                            sS[d][2][ss], float(B.p_ref), float(B.e_ref), F, &F[5]);
        }
        if (task == 2) {
#pragma unroll
          for (int v = 0; v < NFX; ++v) Fin[v] = F[v];
        } else if (d == 1) {
#pragma unroll
          for (int v = 0; v < NFX; ++v) sFj[v][tile_tid(rj, rk)] = F[v];
        } else {
#pragma unroll
          for (int v = 0; v < NFX; ++v) sFk[v][kface_slot(rj, rk)] = F[v];
        }
      }
    }
    // ---- prefetch for plane I+1 (own record of I+2, own centroid of I+1, the ring), issued
    //      before the exchange barrier so that the epilogue and the barrier cover its
    //      latency (this plane's staging consumed the registers it overwrites)
    if (I < Iend) {
      Pn = load_rec32(P, cells, (size_t)(I + 2) * sI + cb);
#pragma unroll
      for (int x = 0; x < 3; ++x) Xn[x] = ldu(XC + x * (size_t)cells, (unsigned)(I + 1) * sI + cb);
      ring_load(I + 2, I + 1);
    }
    if (inner) __syncthreads();                  // (uniform) j / k face exchange
    __builtin_amdgcn_s_setprio(PRIO_REST);
    if (fin) {
      // stages 2 and 3: the base-state loads go out first, so the sums below overlap their
      // latency (before the barrier their registers would spill)
      if (MODE >= 2) {
        const double* qz = q0.p[t.x];
#pragma unroll
        for (int v = 0; v < 5; ++v) vz[v] = ldu(qz + v * cells, c);
      }
      ivol = 1.0 / ivol;
      // outward sums in FP64 of the FP32 face values; the energy gets the gauge terms
      // e_ref x (mass divergence) + p_ref x (sum of w)
      double a[NFX];
#pragma unroll
      for (int v = 0; v < NFX; ++v) a[v] = Cp[v] - (double)Fin[v];
      a[4] = (a[4] + B.e_ref * a[0]) + B.p_ref * a[5];
      double r[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) r[v] = a[v] * ivol;
      if (nd.n_sponge > 0) {
        const double xcv = ldu(XC, c);
        const int i = I - 1 - ng;
        const long long mrow = (long long)i * n1 + j;
        const long long mplane = (long long)n0 * n1;
        for (int sp = 0; sp < nd.n_sponge; ++sp) {
          // a zero weight adds -0 * (q - mean): skipped (rows outside every sponge have no mean)
          const double sigma = sponge_sigma(nd.sponge[sp], xcv);
          if (sigma != 0.0) {
#pragma unroll
            for (int v = 0; v < 5; ++v)
              r[v] = r[v] + (-sigma) * (vq[v] - ldu(B.sponge_mean + v * mplane, (unsigned)mrow));
          }
        }
      }
      double* o = out.p[t.x];
      if (MODE == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) stu(o + v * cells, c, r[v]);
      } else if (MODE == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) stu(o + v * cells, c, vq[v] + dt * r[v]);
      } else {
        const double a0 = MODE == 2 ? 0.75 : 1.0 / 3.0;
        const double a1 = MODE == 2 ? 0.25 : 2.0 / 3.0;
#pragma unroll
        for (int v = 0; v < 5; ++v) stu(o + v * cells, c, a0 * vz[v] + a1 * (vq[v] + dt * r[v]));
      }
    }
    // ---- carry: the partial sum (lower i face minus the j / k faces of this plane's
    //      exchange), gradient record and centroid of the own cell I
    if (act && inner) {
      const int jup = tile_tid(tj + 1, tk), kup = kface_slot(tj, tk + 1);
#pragma unroll
      for (int v = 0; v < NFX; ++v)
        Cp[v] = (double)Fin[v] - (((double)sFj[v][jup] - (double)sFj[v][tid]) +
                                  ((double)sFk[v][kup] - (double)sFk[v][tid]));
    }
    if (own_g) {
#pragma unroll
      for (int v = 0; v < NGW; ++v) Gp[v] = sG[v][sg_own];
#pragma unroll
      for (int x = 0; x < 3; ++x) Xp[x] = sX[x][sg_own];
    }
  }
}

}  // namespace wmx

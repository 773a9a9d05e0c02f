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
#pragma once
#include "flux2.h"

namespace wmx {

namespace {

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

// number of lanes below this one whose bit is set in the wavefront mask m
__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ CellWT<float> rec_w(const Rec32& a) {
  CellWT<float> w;
  w.r = a.r; w.u = a.u; w.v = a.v; w.w = a.w; w.p = a.p; w.e = a.e;
  return w;
}

// Face-vector slots of a tile in LDS: the lower i / j / k face vectors of the cells of the
// gradient region's rows -1 .. 16 (18 rows of pitch G2P; slot gj * G2P + gk), so that the
// upper j / k faces of the gradient cells are the lower faces of the slots one row / column
// on.  They arrive by direct-to-LDS loads (global_load_lds_dword: lane l of a wavefront
// writes slot 64 w' + l of the instruction's run), issued a plane ahead.  NSS covers the
// runs of the last wavefront that loads (360 slots used).
constexpr int NSS_USED = (G2J + 1) * G2P, NSS = 384;
static_assert(NSS_USED <= NSS && NSS % 64 == 0 && NSS - NT <= 128, "face-vector slots");

// Upper i face vector and FP32 volume of a gradient cell (the part of its geometry that is
// not in the face-vector slots).
struct GeoHi {
  float S0hi[3], volr;
};

__device__ __forceinline__ void geo_hi_load(const BlockDev& B, int I, int J, int K, GeoHi& q) {
  const unsigned st0 = (unsigned)B.N[1] * B.N[2];
  const unsigned fc0 = (unsigned)fcount(B.N, 0);
  const unsigned c = (unsigned)I * st0 + (unsigned)J * B.N[2] + (unsigned)K;
#pragma unroll
  for (int x = 0; x < 3; ++x) q.S0hi[x] = ldu(B.Sf[0] + x * (size_t)fc0, c + st0);
  q.volr = ldu((const float*)B.volf, c);
}

// Direct-to-LDS loads of plane I's face-vector slots (slot s -> cell (s / G2P, s % G2P) of
// the region, indices clamped into the block: slots that no gradient cell reads may hold
// any cell's values).  The data lands when the loads complete; the workgroup barrier that
// follows them waits for that (vmcnt) before any wavefront reads the slots.
__device__ __forceinline__ void face_dma(const BlockDev& B, int I, int j0, int k0, float (*sS)[3][NSS], int tid) {
  const int ng = B.ng, N1 = B.N[1], N2 = B.N[2];
  const unsigned fc0 = (unsigned)fcount(B.N, 0), fc1 = (unsigned)fcount(B.N, 1), fc2 = (unsigned)fcount(B.N, 2);
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int s = tid + b * NT;
    if (b == 1 && tid >= NSS - NT) break;                        // (uniform for whole runs)
    const int gj = s / G2P, gk = s - gj * G2P;
    const int J = min(max(ng + j0 - 1 + gj, 0), N1 - 1), K = min(max(ng + k0 - 1 + gk, 0), N2 - 1);
    const unsigned f0 = (unsigned)I * ((unsigned)N1 * N2) + (unsigned)J * N2 + (unsigned)K;
    const unsigned f1 = (unsigned)I * ((unsigned)(N1 + 1) * N2) + (unsigned)J * N2 + (unsigned)K;
    const unsigned f2 = (unsigned)I * ((unsigned)N1 * (N2 + 1)) + (unsigned)J * (unsigned)(N2 + 1) + (unsigned)K;
    const int run = s & ~63;                                     // the wavefront's 64-slot run
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      __builtin_amdgcn_global_load_lds((const void*)(B.Sf[0] + x * (size_t)fc0 + f0),
                                       (__attribute__((address_space(3))) void*)&sS[0][x][run], 4, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(B.Sf[1] + x * (size_t)fc1 + f1),
                                       (__attribute__((address_space(3))) void*)&sS[1][x][run], 4, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(B.Sf[2] + x * (size_t)fc2 + f2),
                                       (__attribute__((address_space(3))) void*)&sS[2][x][run], 4, 0, 0);
    }
  }
}

// Gradient record of a cell (P-region slot ps, face-vector slot gsl) from the staged planes,
// the face-vector slots and q: the per-cell arithmetic of k_grad32 (grad32.h), faces in the
// order i, j, k.
__device__ __forceinline__ void grad_cell3(const NumDev& nd, const GeoHi& q, const float (*sS)[3][NSS], int gsl,
                                           const float (*pm)[NP2], const float (*p0)[NP2], const float (*pp)[NP2],
                                           const float (*pr)[NP2], int ps, float* G) {
  float S0lo[3], S0hi[3], S1lo[3], S1hi[3], S2lo[3], S2hi[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    S0lo[x] = sS[0][x][gsl];
    S0hi[x] = q.S0hi[x];
    S1lo[x] = sS[1][x][gsl];
    S1hi[x] = sS[1][x][gsl + G2P];
    S2lo[x] = sS[2][x][gsl];
    S2hi[x] = sS[2][x][gsl + 1];
  }
  const float volr = q.volr;
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
  __shared__ float sS[3][3][NSS];                // face-vector slots (lower i / j / k faces)
  __shared__ float sFj[NFX][NT];                 // slot tile_tid(r, c)
  __shared__ float sFk[NFX][NT + C2J + 1];       // slot kface_slot(r, c)
  __shared__ unsigned short sUq[NT / 64][2 * 64];  // per wavefront: j / k faces with upwind work
  const int4 t = tiles[tile];
  const BlockDev& B = blks[t.x];
  const int j0 = t.y, k0 = t.z, i0 = t.w;
  const int i1 = min(i0 + ci_len, B.n[0]);
  const int tid = threadIdx.x;
  const int tj = (tid >> 6) + 4 * ((tid >> 4) & 3), tk = tid & 15;   // tile_tid(tj, tk) == tid
  const int wv = tid >> 6, lane = tid & 63;
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
  // first gradient cell of each thread: region cell tid at owned planes, the own cell
  // (tj + 1, tk + 1) at the two end planes; indices clamped into the block
  auto cell_of = [&](int r, int& gj, int& gk, int& JJ, int& KK) {
    gj = r / G2K; gk = r - gj * G2K;
    JJ = ng + j0 - 1 + gj; KK = ng + k0 - 1 + gk;
    const bool ok = g_ok(JJ, KK);
    JJ = min(max(JJ, 0), N[1] - 1); KK = min(max(KK, 0), N[2] - 1);
    return ok;
  };
  auto first_cell = [&](bool at_inner) { return at_inner ? tid : (tj + 1) * G2K + (tk + 1); };
  GeoHi qn;                                      // upper i face and volume of the next plane's first cell
  auto next_geometry = [&](int P) {              // plane P's face-vector slots and qn
    face_dma(B, P, j0, k0, sS, tid);
    int gj, gk, JJ, KK;
    cell_of(first_cell(P >= I0 && P < Iend), gj, gk, JJ, KK);
    geo_hi_load(B, P, JJ, KK, qn);
  };
  next_geometry(I0 - 1);
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
    // this plane's face-vector slots were loaded straight into LDS during the previous
    // plane: every wavefront waits for its vector memory operations to complete (vmcnt(0);
    // the barrier itself waits for LDS operations only) before the staging barrier
    __builtin_amdgcn_s_waitcnt(WAITCNT_VM0);
    __syncthreads();
    // ---- gradient records of plane I: the tile + ring (306 cells) at owned planes, the own
    //      cells at the two end planes.  The first cell's upper i face and volume were loaded
    //      during the previous plane; wavefront 0 also takes the region's last 50 cells.
    //      Cells outside the block's cells and first ghost layer are evaluated on clamped
    //      indices and not stored.
    {
      auto eval_store = [&](const GeoHi& q, int gj, int gk, bool ok) {
        float Gv[NGW];
        const int gsl = gj * G2P + gk;
        grad_cell3(nd, q, sS, gsl, sPh[s_m], sPh[s_0], sPh[s_p], sPr, (gj + 1) * H2K + (gk + 1), Gv);
        if (ok) {
#pragma unroll
          for (int v = 0; v < NGW; ++v) sG[v][gsl] = Gv[v];
        }
      };
      if (inner || own) {
        int gj0, gk0, JJ0, KK0;
        const bool ok0 = cell_of(first_cell(inner), gj0, gk0, JJ0, KK0);
        eval_store(qn, gj0, gk0, ok0);
      }
      if (inner && tid + NT < NG2) {
        int gj1, gk1, JJ1, KK1;
        const bool ok1 = cell_of(tid + NT, gj1, gk1, JJ1, KK1);
        GeoHi q1;
        geo_hi_load(B, I, JJ1, KK1, q1);
        eval_store(q1, gj1, gk1, ok1);
      }
    }
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
    // Hybrid configurations: the tasks evaluate the central and viscous parts; the upwind
    // (MUSCL-HLLC) parts of the faces with s > 0 are evaluated afterwards in packed passes
    // of the wavefront (the j / k faces listed in sUq, the own i face flagged in up_i), so
    // that a wavefront runs the upwind arithmetic once per plane instead of once per task
    constexpr bool DEFER = SC == SC_HYB_KEP4_ALBADA || SC == SC_ANY;
    int nq = 0;                                  // listed faces of the wavefront (uniform)
    bool up_i = false;
#pragma unroll 1
    for (int task = inner ? 0 : 2; task < (I >= I0 ? 3 : 0); ++task) {
      bool doit = act, later = false;
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
          const float sw = flux_core_f32<SC, DEFER>(nd, L2, L, Rc, R2, thL, thR, gL, gR, TL, TR, dx[0], dx[1], dx[2],
                                                    sS[d][0][ss], sS[d][1][ss], sS[d][2][ss], float(B.p_ref),
                                                    float(B.e_ref), F, &F[5]);
          later = DEFER && sw > 0.f;
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
      if (DEFER) {
        if (task == 2) {
          up_i = later;
        } else {
          const unsigned long long m = __ballot(later);
          if (later) sUq[wv][nq + lanes_below(m)] = (unsigned short)((d << 10) | (rj << 5) | rk);
          nq += __popcll(m);
        }
      }
    }
    if (DEFER) {
      // packed upwind passes: in the first, the lanes whose i face has upwind work take it
      // and the other lanes the first listed faces; then 64 listed faces per pass
      const unsigned long long mi = __ballot(up_i);
      const int ni = __popcll(mi), rank_free = lanes_below(~mi);
      const float pr = float(B.p_ref), er = float(B.e_ref);
      for (int qb = 0, pass = 0; (pass == 0 && ni > 0) || qb < nq; ++pass) {
        const bool mine = pass == 0 && up_i;
        const int q = qb + (pass == 0 ? rank_free : lane);
        const bool lst = !mine && q < nq;
        if (mine || lst) {
          CellWT<float> L2, L, Rc, R2;
          float thL, thR, sv[3];
          int dd = 0, rj = 0, rk = 0;
          if (mine) {
            L2 = rec_w(Pw[0]); L = rec_w(Pw[1]); Rc = rec_w(Pw[2]); R2 = rec_w(Pw[3]);
            thL = Gp[14]; thR = sG[14][sg_own];
#pragma unroll
            for (int x = 0; x < 3; ++x) sv[x] = sS[0][x][sg_own];
          } else {
            const int dsc = sUq[wv][q];
            dd = dsc >> 10; rj = (dsc >> 5) & 31; rk = dsc & 31;
            const int pR = (rj + 2) * H2K + (rk + 2), gR_ = (rj + 1) * G2P + (rk + 1);
            const int pst = dd == 1 ? H2K : 1, gst = dd == 1 ? G2P : 1;
            auto cw = [&](int s) {
              CellWT<float> w;
              w.r = sPr[0][s]; w.u = sPh[s_0][0][s]; w.v = sPh[s_0][1][s]; w.w = sPh[s_0][2][s];
              w.p = sPr[1][s]; w.e = sPr[3][s];
              return w;
            };
            L2 = cw(pR - 2 * pst); L = cw(pR - pst); Rc = cw(pR); R2 = cw(pR + pst);
            thL = sG[14][gR_ - gst]; thR = sG[14][gR_];
#pragma unroll
            for (int x = 0; x < 3; ++x) sv[x] = sS[dd][x][gR_];
          }
          float dF[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
          upwind_add<SC>(nd, face_weight<SC>(nd, thL, thR), L2, L, Rc, R2, sv[0], sv[1], sv[2], pr, er, dF);
          if (mine) {
#pragma unroll
            for (int v = 0; v < 5; ++v) Fin[v] = Fin[v] + dF[v];
          } else if (dd == 1) {
            const int sl = tile_tid(rj, rk);
#pragma unroll
            for (int v = 0; v < 5; ++v) sFj[v][sl] = sFj[v][sl] + dF[v];
          } else {
            const int sl = kface_slot(rj, rk);
#pragma unroll
            for (int v = 0; v < 5; ++v) sFk[v][sl] = sFk[v][sl] + dF[v];
          }
        }
        qb += pass == 0 ? 64 - ni : 64;
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
    // the face-vector slots are free (this plane's face tasks are done): the next plane's
    // go out now, and the epilogue covers their latency
    if (I < Iend) next_geometry(I + 1);
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

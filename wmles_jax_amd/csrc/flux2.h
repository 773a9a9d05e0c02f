// FP32-bulk fused face fluxes + divergence + SSP-RK3 stage update of one marching tile
// (flux2_tile), shared by the k_flux2 kernel (flux.hip) and the persistent stage kernel
// (stage_fast.hip); the unit that includes it fixes the FP contraction (fast).
//
// One workgroup owns a 15 x 16 (j, k) column tile and marches i over a chunk of planes.
// Each thread keeps the i-direction stencil of its own column in registers (primitives
// at i-1..i+2, gradient records and centroids at i, i+1), so every plane is read from
// HBM once.  The current plane is staged in LDS with the halo the j/k faces need
// (primitives: 2 cells, gradient records and centroids: 1 cell).  Every face flux is
// evaluated once: the i face of each column in registers, the j/k faces of the tile
// (plus its upper edge rows) through LDS, and each cell sums its six faces (reference
// solver.py:86-95: inviscid scheme + viscous flux, wall override, divergence / vol, then
// sources, solver.py:103-104, and the stage combination, solver.py:109-121).
#pragma once
#include "common.h"

namespace wmx {

namespace {

constexpr int TJ = 16, TK = 16, NT = TJ * TK;
constexpr int HJ = TJ + 4, HK = TK + 4;      // primitive region (halo 2)
constexpr int GJ = TJ + 2, GK = TK + 2;      // gradient / centroid region (halo 1)
constexpr int NPW = 7;                       // r u v w p e T
constexpr int NGW = 15;                      // g[12], mu_eff, k_eff, theta
constexpr int NFX = 6;                       // FP32-bulk face exchange: F'[5], w

template <class R>
struct ColT {                                // one cell's primitive record + T
  CellWT<R> w;
  R T;
};

template <class R>
struct ColGT {
  CellGT<R> g;
  R th;
};

template <class R, class PT>
__device__ __forceinline__ ColT<R> load_col(PT P, long long cells, long long c) {
  ColT<R> o;
  o.w.r = P[c];
  o.w.u = P[cells + c];
  o.w.v = P[2 * cells + c];
  o.w.w = P[3 * cells + c];
  o.w.p = P[4 * cells + c];
  o.w.e = P[P_E * cells + c];
  o.T = P[P_T * cells + c];
  return o;
}

// the same from the FP32-bulk records (two / four 16-byte loads)
__device__ __forceinline__ ColT<float> load_col32(const float* P, size_t cells, size_t c) {
  const float4 a = ld4(P + 4 * c), b = ld4(P + 4 * (cells + c));
  ColT<float> o;
  o.w.u = a.x; o.w.v = a.y; o.w.w = a.z; o.T = a.w;
  o.w.r = b.x; o.w.p = b.y; o.w.e = b.w;
  return o;
}

__device__ __forceinline__ ColGT<float> load_colg32(const float* G, size_t cells, size_t c) {
  const float* r = G + 4 * c;
  const size_t q = 4 * cells;
  const float4 a = ld4(r), b = ld4(r + q), d = ld4(r + 2 * q), e = ld4(r + 3 * q);
  ColGT<float> o;
  o.g.g[0] = a.x; o.g.g[1] = a.y; o.g.g[2] = a.z; o.g.g[3] = a.w;
  o.g.g[4] = b.x; o.g.g[5] = b.y; o.g.g[6] = b.z; o.g.g[7] = b.w;
  o.g.g[8] = d.x; o.g.g[9] = d.y; o.g.g[10] = d.z; o.g.g[11] = d.w;
  o.g.mu = e.x; o.g.k = e.y; o.th = e.z;
  return o;
}

template <class R, class PT>
__device__ __forceinline__ ColGT<R> load_colg(PT G, long long cells, long long c) {
  ColGT<R> o;
#pragma unroll
  for (int i = 0; i < 12; ++i) o.g.g[i] = G[i * cells + c];
  o.g.mu = G[G_MU * cells + c];
  o.g.k = G[G_K * cells + c];
  o.th = G[G_TH * cells + c];
  return o;
}

// FP32-bulk face flux in gauge form.  The records carry p' = p - p_ref and
// e' = e_int - e_ref.  The full flux through face f is
//   F_f = (1 - s) (Fc'_f + p_ref S_f [momentum] + e_ref m_f + p_ref u_f [energy]) + s Fu_f
// (Fc' the central flux of the shifted records, m_f its mass flux, u_f its normal
// velocity times the area, Fu the MUSCL-HLLC flux with the full pressure, s the hybrid
// weight) plus the viscous flux.  Two identities move every large term out of the
// per-face FP32 values:
//  * sum over a cell's faces of p_ref S_f (outward) is p_ref times the closure of the
//    FP64 geometry (zero to FP64 rounding), so (1 - s) p_ref S_f is replaced by
//    -s p_ref S_f, which is non-zero only on upwind-blended faces;
//  * (1 - s) m_f = M_f - s Fu_f[0], with M_f the blended mass flux the face already
//    carries, so e_ref (1 - s) m_f sums to e_ref times the cell's mass divergence, which
//    the FP64 epilogue forms from the same FP32 values.
// The face therefore hands over F'_f = (1 - s) Fc'_f + s Fu_f (momentum - s p_ref S_f,
// energy - s e_ref Fu_f[0]) + viscous, and w_f = (1 - s) u_f.  The epilogue adds
// e_ref x (mass divergence) + p_ref x (sum of w_f) to the energy in FP64.  On faces with
// s = 0 (central) every term keeps FP64 accuracy; on blended faces the error is that of
// the FP32 HLLC flux itself.
__device__ __forceinline__ float hybrid_weight(const NumDev& nd, float thL, float thR) {
  const float th = fmaxf(thL, thR);
  return fminf(fmaxf(fdiv<float>(th - float(nd.th0), float(nd.th1 - nd.th0)), 0.f), 1.f);
}

// Scheme configurations compiled as their own kernel instances (the face flux then
// carries no runtime dispatch); SC_ANY reads everything from NumDev.
// (the configurations SC_* and scheme_config are in common.h)

// Kennedy-Gruber central flux of order 2 or 4 in FP32 (reference schemes.py:15-23,57-70),
// the same quantities as central_flux<false, float> with fewer operations: each cell's
// e + |u|^2 / 2 is formed once for the face's stencil, and the three two-point fluxes are
// kept as pair sums (sr = rho_a + rho_b, su, sp, sE; w = su . S, M = sr w), whose scale
// factors (1/4, 1/8, 1/2) fold into the order-4 combination:
//   F = (8 A_LR - (A_LR2 + A_L2R)) / 48 per momentum / energy component with
//   A_x = M su_x + 4 sp S_x, A_4 = M sE + 2 sp w; mass (8 M_LR - ...) / 24; u.S (8 w - ...) / 12.
template <int ORDER>
__device__ __forceinline__ void kep_pairs_f32(const CellWT<float>& L2, const CellWT<float>& L,
                                              const CellWT<float>& Rc, const CellWT<float>& R2, float Sx,
                                              float Sy, float Sz, float* F, float* un4) {
  auto ek = [](const CellWT<float>& c) { return c.e + 0.5f * ((c.u * c.u + c.v * c.v) + c.w * c.w); };
  const float EL = ek(L), ER = ek(Rc);
  auto pair = [&](const CellWT<float>& a, const CellWT<float>& b, float Ea, float Eb, float* A, float* w) {
    const float sr = a.r + b.r, su = a.u + b.u, sv = a.v + b.v, sw = a.w + b.w;
    const float sp = a.p + b.p, sE = Ea + Eb;
    const float ww = (su * Sx + sv * Sy) + sw * Sz;
    const float M = sr * ww, sp4 = 4.f * sp;
    A[0] = M;
    A[1] = M * su + sp4 * Sx;
    A[2] = M * sv + sp4 * Sy;
    A[3] = M * sw + sp4 * Sz;
    A[4] = M * sE + (2.f * sp) * ww;
    *w = ww;
  };
  float A[5], w;
  pair(L, Rc, EL, ER, A, &w);
  if (ORDER == 2) {
    F[0] = 0.25f * A[0];
#pragma unroll
    for (int c = 1; c < 5; ++c) F[c] = 0.125f * A[c];
    *un4 = 0.5f * w;
    return;
  }
  float Aa[5], wa, Ab[5], wb;
  pair(L, R2, EL, ek(R2), Aa, &wa);
  pair(L2, Rc, ek(L2), ER, Ab, &wb);
  F[0] = (8.f * A[0] - (Aa[0] + Ab[0])) * (1.f / 24.f);
#pragma unroll
  for (int c = 1; c < 5; ++c) F[c] = (8.f * A[c] - (Aa[c] + Ab[c])) * (1.f / 48.f);
  *un4 = (8.f * w - (wa + wb)) * (1.f / 12.f);
}

template <int SC>
__device__ __forceinline__ void flux_core_f32(const NumDev& nd, const CellWT<float>& L2,
                                              const CellWT<float>& L, const CellWT<float>& Rc,
                                              const CellWT<float>& R2, float thL, float thR,
                                              const CellGT<float>& gL, const CellGT<float>& gR,
                                              float TL, float TR, float d0, float d1, float d2,
                                              float sx, float sy, float sz, float pr, float er,
                                              float* Ff, float* uw) {
  const int sch = SC == SC_HYB_KEP4_ALBADA ? WMX_SCHEME_HYBRID
                  : (SC == SC_KEP2 || SC == SC_KEP4) ? WMX_SCHEME_KEP : nd.scheme;
  const int order = SC == SC_KEP2 ? 2 : ((SC == SC_HYB_KEP4_ALBADA || SC == SC_KEP4) ? 4 : nd.order);
  const int limiter = SC == SC_HYB_KEP4_ALBADA ? 0 : nd.limiter;
  float s = sch == WMX_SCHEME_MUSCLHLLC ? 1.f : 0.f;
  if (sch == WMX_SCHEME_HYBRID) s = hybrid_weight(nd, thL, thR);
  float Fc[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, Fu[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, un4 = 0.f;
  if (s < 1.f) {
    const bool keep = SC == SC_ANY && (sch == WMX_SCHEME_KEEP ||
                                       (sch == WMX_SCHEME_HYBRID && nd.central == WMX_SCHEME_KEEP));
    if (keep) central_flux<true, float>(order, L2, L, Rc, R2, sx, sy, sz, Fc, &un4);
    else if (order == 2) kep_pairs_f32<2>(L2, L, Rc, R2, sx, sy, sz, Fc, &un4);
    else kep_pairs_f32<4>(L2, L, Rc, R2, sx, sy, sz, Fc, &un4);
  }
  const float wc = 1.f - s;
#pragma unroll
  for (int v = 0; v < 5; ++v) Ff[v] = wc * Fc[v];
  if (s > 0.f) {
    muscl_hllc_flux<float>(float(nd.gas.gamma), limiter, L2, L, Rc, R2, sx, sy, sz, pr, Fu);
    Ff[0] = Ff[0] + s * Fu[0];
    Ff[1] = Ff[1] + s * (Fu[1] - pr * sx);
    Ff[2] = Ff[2] + s * (Fu[2] - pr * sy);
    Ff[3] = Ff[3] + s * (Fu[3] - pr * sz);
    Ff[4] = Ff[4] + s * (Fu[4] - er * Fu[0]);
  }
  *uw = wc * un4;
  const float dist = fsqrt<float>((d0 * d0 + d1 * d1) + d2 * d2);
  const float id = fdiv<float>(1.f, dist);
  const float phL[4] = {L.u, L.v, L.w, TL};
  const float phR[4] = {Rc.u, Rc.v, Rc.w, TR};
  viscous_face_e<float>(gL, gR, phL, phR, d0 * id, d1 * id, d2 * id, id, sx, sy, sz, Ff);
}

template <class R>
__device__ __forceinline__ void flux_core(const NumDev& nd, const CellWT<R>& L2, const CellWT<R>& L,
                                          const CellWT<R>& Rc, const CellWT<R>& R2, R thL, R thR,
                                          const CellGT<R>& gL, const CellGT<R>& gR, R TL, R TR,
                                          const double* xL, const double* xR, double Sx, double Sy,
                                          double Sz, double* F) {
  static_assert(sizeof(R) == 8, "the FP32-bulk face flux is flux_core_f32");
  const R sx = R(Sx), sy = R(Sy), sz = R(Sz);
  const int sch = nd.scheme;
  double Fc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Fu[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double s = sch == WMX_SCHEME_MUSCLHLLC ? 1.0 : 0.0;
  if (sch == WMX_SCHEME_HYBRID) {
    const R th = fmax(thL, thR);
    s = (double)fmin(fmax((th - R(nd.th0)) / R(nd.th1 - nd.th0), R(0.0)), R(1.0));
  }
  if (s < 1.0) {
    R Fr[5], un4;
    const bool keep = sch == WMX_SCHEME_KEEP || (sch == WMX_SCHEME_HYBRID && nd.central == WMX_SCHEME_KEEP);
    if (keep) central_flux<true, R>(nd.order, L2, L, Rc, R2, sx, sy, sz, Fr, &un4);
    else central_flux<false, R>(nd.order, L2, L, Rc, R2, sx, sy, sz, Fr, &un4);
#pragma unroll
    for (int v = 0; v < 5; ++v) Fc[v] = (double)Fr[v];
  }
  if (s > 0.0) {
    R Fr[5];
    muscl_hllc_flux<R>(R(nd.gas.gamma), nd.limiter, L2, L, Rc, R2, sx, sy, sz, R(0), Fr);
#pragma unroll
    for (int v = 0; v < 5; ++v) Fu[v] = (double)Fr[v];
  }
  if (sch == WMX_SCHEME_HYBRID) {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = (1.0 - s) * Fc[v] + s * Fu[v];
  } else if (sch == WMX_SCHEME_MUSCLHLLC) {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = Fu[v];
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = Fc[v];
  }
  const double dx0 = xR[0] - xL[0], dx1 = xR[1] - xL[1], dx2 = xR[2] - xL[2];
  const double idist = 1.0 / sqrt((dx0 * dx0 + dx1 * dx1) + dx2 * dx2);
  const R phL[4] = {L.u, L.v, L.w, TL};
  const R phR[4] = {Rc.u, Rc.v, Rc.w, TR};
  R Fv[5] = {R(0), R(0), R(0), R(0), R(0)};
  viscous_face_e<R>(gL, gR, phL, phR, R(dx0 * idist), R(dx1 * idist), R(dx2 * idist), R(idist), sx, sy,
                    sz, Fv);
#pragma unroll
  for (int v = 0; v < 5; ++v) F[v] = F[v] + (double)Fv[v];
}

// FP32-bulk tile: 15 rows (j) x 16 columns (k) of cells on a 16 x 16 thread grid.
// Per plane the tile has 240 i faces, 16 x 16 j faces (incl. the upper edge row) and
// 15 x 17 k faces (incl. the upper edge column): thread (tj, tk) evaluates the i face
// of its cell (tj < 15), the j face below row tj, and the k face left of its cell
// (tj < 15) or the upper edge k face of row tk (tj = 15) -- three face tasks for every
// wavefront, none waiting at the plane barrier for a fourth.
constexpr int C2J = FLUX2_TILE_J;              // cell rows of an FP32-bulk tile
constexpr int H2J = C2J + 4, H2K = TK + 4;     // primitive region (halo 2)
constexpr int G2J = C2J + 2, G2K = TK + 2;     // gradient / centroid region (halo 1)
constexpr int NP2 = H2J * H2K, NG2 = G2J * G2K;
// LDS banks: a wavefront covers tile rows w, w+4, w+8, w+12 (w = wavefront index) with
// 16 columns each.  Rows of pitch 20 then start 80 words apart, 16 banks apart modulo 32
// (and 16, 32, 48 modulo 64), so every row-of-16 access of a wavefront is conflict-free.
// The gradient / centroid region is stored with that pitch too (18 columns used).
constexpr int G2P = 20, NG2S = G2J * G2P;
static_assert(H2K == 20 && G2P >= G2K && TK == 16 && NT == 256, "bank-free layout assumes 16 x 16");
// face exchange slots: a face with tile coordinates (r, c), c < 16, sits at the index of
// the thread that owns (r, c); the upper edge k faces (c = 16) follow at 256 + r
__device__ __forceinline__ int tile_tid(int r, int c) { return (r & 3) * 64 + (r >> 2) * 16 + c; }
__device__ __forceinline__ int kface_slot(int r, int c) { return c < TK ? tile_tid(r, c) : NT + r; }

// slot t of the halo ring of width h around a rows x TK interior
__device__ __forceinline__ void ring_slot(int t, int h, int rows, int* sj, int* sk) {
  const int W = TK + 2 * h, top = h * W;
  if (t < top) { *sj = t / W; *sk = t % W; return; }
  if (t < 2 * top) { t -= top; *sj = rows + h + t / W; *sk = t % W; return; }
  t -= 2 * top;
  const int row = t / (2 * h), c = t % (2 * h);
  *sj = h + row;
  *sk = c < h ? c : TK + c;
}

template <class R>
__device__ __forceinline__ CellWT<R> sp_w(const R (*sP)[NP2], int s) {
  CellWT<R> w;
  w.r = sP[0][s]; w.u = sP[1][s]; w.v = sP[2][s]; w.w = sP[3][s]; w.p = sP[4][s]; w.e = sP[5][s];
  return w;
}

template <class R>
__device__ __forceinline__ CellGT<R> sg_g(const R (*sG)[NG2S], int s) {
  CellGT<R> g;
#pragma unroll
  for (int i = 0; i < 12; ++i) g.g[i] = sG[i][s];
  g.mu = sG[12][s];
  g.k = sG[13][s];
  return g;
}

}  // namespace

// One marching tile (tiles[tile]) of the FP32-bulk flux launch; called by every thread
// of a 256-thread workgroup (the k_flux2 kernel, and the persistent stage kernel for each
// tile it takes).  MODE 0: r = RHS; 1: q + dt r; 2: 3/4 q0 + 1/4 (q + dt r); 3: 1/3 q0 +
// 2/3 (q + dt r)
template <class R, int MODE, int SC>
__device__ __forceinline__ void flux2_tile(const BlockDev* __restrict__ blks,
                                           const int4* __restrict__ tiles, int ci_len,
                                           const NumDev& nd, const Ptrs& qin, const Ptrs& q0,
                                           const Ptrs& out, const double* __restrict__ d_dt,
                                           double dt_host, int tile) {
  // staging area: primitives (halo 2) and gradient records (halo 1) in R; the FP64
  // face fluxes of the plane go to their own exchange buffers as they are computed
  __shared__ R sP[NPW][NP2];
  __shared__ R sG[NGW][NG2S];
  __shared__ double sX[3][NG2S];
  // face exchange: F'[5] and w (gauge form, FP32)
  __shared__ float sFj[NFX][NT];                 // slot tile_tid(r, c)
  __shared__ float sFk[NFX][NT + C2J + 1];       // slot kface_slot(r, c)
  __shared__ float sFi[2][NFX][C2J * TK];        // own i faces by thread: lower / upper, ping-pong
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
  const bool own = tj < C2J;                     // thread owns a column of the tile
  const bool act = own && j < n1 && k < n2;
  const int jl = min(j0 + C2J, n1) - j0;
  const int kl = min(k0 + TK, n2) - k0;
  const int J = ng + j, K = ng + k;
  // 32-bit element offsets from wave-uniform plane bases (every plane < 4 GiB)
  const unsigned cells = (unsigned)B.cells;
  const unsigned sI = (unsigned)N[1] * N[2];
  const double dt = d_dt ? d_dt[B.case_id] : dt_host;
  static_assert(sizeof(R) == 4, "k_flux2 is the FP32-bulk kernel");
  const R* P = (const R*)B.Pf;
  const R* G = (const R*)B.Gf;
  const double* XC = B.xc;
  // boxes the staged data may come from: primitives [0, N), gradients [ng-1, ng+n]
  auto p_ok = [&](int jj, int kk) { return jj >= 0 && jj < N[1] && kk >= 0 && kk < N[2]; };
  auto g_ok = [&](int jj, int kk) {
    return jj >= ng - 1 && jj <= ng + n1 && kk >= ng - 1 && kk <= ng + n2;
  };
  // i-direction register window of the own column
  // The window is loaded by every thread (idle ones from a clamped, valid column), so
  // each load lands in its window register directly and stays in flight until used.
  ColT<R> Pw[4];
  ColGT<R> Gw[2];
  double Xw[2][3];
  const int I0 = ng + i0;
  const unsigned cb = (unsigned)min(J, N[1] - 1) * N[2] + min(K, N[2] - 1);
#pragma unroll
  for (int s = 1; s < 4; ++s) Pw[s] = load_col32(P, cells, (size_t)(I0 - 3 + s) * sI + cb);
  Gw[1] = load_colg32(G, cells, (size_t)(I0 - 1) * sI + cb);
#pragma unroll
  for (int x = 0; x < 3; ++x) Xw[1][x] = ldu(XC + x * cells, (unsigned)(I0 - 1) * sI + cb);
  const int own_s = tid;                        // this column's i-face slot (own threads: tid < 240)
  int lo = 0;                                    // slot of the current lower i face
  // halo ring job of this thread (plane-independent): 140 primitive slots (halo 2),
  // 66 gradient/centroid slots (halo 1); prefetched one plane ahead into registers
  constexpr int NRP = NP2 - C2J * TK, NRG = NG2 - C2J * TK;
  int ring = 0, ring_s = 0;                      // 1: primitive slot, 2: gradient slot
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
  R rv[NGW];
  double rx[3];
  auto ring_load = [&](int Ip) {
    const unsigned c = (unsigned)Ip * sI + ring_off;
    if (ring == 1) {
      const ColT<float> o = load_col32(P, cells, c);
      rv[0] = o.w.r; rv[1] = o.w.u; rv[2] = o.w.v; rv[3] = o.w.w; rv[4] = o.w.p; rv[5] = o.w.e; rv[6] = o.T;
    } else if (ring == 2) {
      const ColGT<float> o = load_colg32(G, cells, c);
#pragma unroll
      for (int i = 0; i < 12; ++i) rv[i] = o.g.g[i];
      rv[12] = o.g.mu; rv[13] = o.g.k; rv[14] = o.th;
#pragma unroll
      for (int x = 0; x < 3; ++x) rx[x] = ldu(XC + x * cells, c);
    }
  };

  for (int I = I0 - 1; I < ng + i1; ++I) {
    const bool first = I == I0 - 1;
    // ---- shift the window (planes I-1..I+1 of primitives, I of gradients/centroids)
    Pw[0] = Pw[1]; Pw[1] = Pw[2]; Pw[2] = Pw[3];
    Gw[0] = Gw[1];
#pragma unroll
    for (int x = 0; x < 3; ++x) Xw[0][x] = Xw[1][x];
    // ---- stage plane I in LDS: own columns from registers, halo and idle columns from
    //      HBM (each group of loads is issued as a batch before its LDS stores)
    if (!first) {
      const int sp = (tj + 2) * H2K + (tk + 2), sg = (tj + 1) * G2P + (tk + 1);
      const unsigned pl = (unsigned)I * sI;
      if (act) {
        sP[0][sp] = Pw[1].w.r; sP[1][sp] = Pw[1].w.u; sP[2][sp] = Pw[1].w.v; sP[3][sp] = Pw[1].w.w;
        sP[4][sp] = Pw[1].w.p; sP[5][sp] = Pw[1].w.e; sP[6][sp] = Pw[1].T;
#pragma unroll
        for (int i = 0; i < 12; ++i) sG[i][sg] = Gw[0].g.g[i];
        sG[12][sg] = Gw[0].g.mu; sG[13][sg] = Gw[0].g.k; sG[14][sg] = Gw[0].th;
#pragma unroll
        for (int x = 0; x < 3; ++x) sX[x][sg] = Xw[0][x];
      } else if (own) {                              // idle column of a partial tile
        const unsigned c = pl + (unsigned)J * N[2] + K;
        if (p_ok(J, K)) {
          const ColT<float> o = load_col32(P, cells, c);
          const R vp[NPW] = {o.w.r, o.w.u, o.w.v, o.w.w, o.w.p, o.w.e, o.T};
#pragma unroll
          for (int i = 0; i < NPW; ++i) sP[i][sp] = vp[i];
        }
        if (g_ok(J, K)) {
          const ColGT<float> og = load_colg32(G, cells, c);
          R vg[NGW];
          double vx[3];
#pragma unroll
          for (int i = 0; i < 12; ++i) vg[i] = og.g.g[i];
          vg[12] = og.g.mu; vg[13] = og.g.k; vg[14] = og.th;
#pragma unroll
          for (int x = 0; x < 3; ++x) vx[x] = ldu(XC + x * cells, c);
#pragma unroll
          for (int i = 0; i < NGW; ++i) sG[i][sg] = vg[i];
#pragma unroll
          for (int x = 0; x < 3; ++x) sX[x][sg] = vx[x];
        }
      }
      // halo ring from the registers prefetched during the previous plane
      if (ring == 1) {
#pragma unroll
        for (int i = 0; i < NPW; ++i) sP[i][ring_s] = rv[i];
      } else if (ring == 2) {
#pragma unroll
        for (int i = 0; i < NGW; ++i) sG[i][ring_s] = rv[i];
#pragma unroll
        for (int x = 0; x < 3; ++x) sX[x][ring_s] = rx[x];
      }
    }
    // ---- load plane I+2 (primitives) and I+1 (gradients, centroids) into the window;
    //      issued after the halo loads (completion counts in issue order), needed by
    //      the last face task
    Pw[3] = load_col32(P, cells, (size_t)(I + 2) * sI + cb);
    Gw[1] = load_colg32(G, cells, (size_t)(I + 1) * sI + cb);
#pragma unroll
    for (int x = 0; x < 3; ++x) Xw[1][x] = ldu(XC + x * cells, (unsigned)(I + 1) * sI + cb);
    __syncthreads();
    // ---- face fluxes: 0 j face below row tj, 1 k face left of column tk (rows tj < 15)
    //      or the upper edge k face of row tk (tj = 15), 2 i face I+1 (registers)
#pragma unroll 1
    for (int task = first ? 2 : 0; task < 3; ++task) {
      bool doit = act;
      int d = 0, fI = I, fJ = J, fK = K;
      int rj = tj, rk = tk;                          // tile coords of the face's right cell
      CellWT<R> L2, L, Rc, R2;
      CellGT<R> gL, gR;
      R thL, thR, TL, TR;
      float dx[3] = {0.f, 0.f, 0.f};           // centroid difference, FP64 then rounded
      if (task == 2) {
        fI = I + 1;
        L2 = Pw[0].w; L = Pw[1].w; Rc = Pw[2].w; R2 = Pw[3].w;
        TL = Pw[1].T; TR = Pw[2].T;
        gR = Gw[1].g; thR = Gw[1].th;
        if (first) {
          gL = Gw[0].g; thL = Gw[0].th;
#pragma unroll
          for (int x = 0; x < 3; ++x) dx[x] = float(Xw[1][x] - Xw[0][x]);
        } else {                                     // plane I of the own column, staged
          const int sg = (tj + 1) * G2P + (tk + 1);
          gL = sg_g<R>(sG, sg); thL = sG[14][sg];
#pragma unroll
          for (int x = 0; x < 3; ++x) dx[x] = float(Xw[1][x] - sX[x][sg]);
        }
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
          L2 = sp_w<R>(sP, pR - 2 * pst); L = sp_w<R>(sP, pR - pst); Rc = sp_w<R>(sP, pR);
          R2 = sp_w<R>(sP, pR + pst);
          TL = sP[6][pR - pst]; TR = sP[6][pR];
          gL = sg_g<R>(sG, gR_ - gst); gR = sg_g<R>(sG, gR_);
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
          const unsigned fi = (unsigned)fidx(N, d, fI, fJ, fK), fc = (unsigned)fcount(N, d);
          const float* S = B.Sf[d];
          flux_core_f32<SC>(nd, L2, L, Rc, R2, thL, thR, gL, gR, TL, TR, dx[0], dx[1], dx[2], ldu(S, fi),
                            ldu(S + fc, fi), ldu(S + 2 * fc, fi), float(B.p_ref), float(B.e_ref), F,
                            &F[5]);
        }
        if (task == 2) {
          const int hs = first ? 0 : lo ^ 1;
#pragma unroll
          for (int v = 0; v < NFX; ++v) sFi[hs][v][own_s] = F[v];
        } else if (d == 1) {
#pragma unroll
          for (int v = 0; v < NFX; ++v) sFj[v][tile_tid(rj, rk)] = F[v];
        } else {
#pragma unroll
          for (int v = 0; v < NFX; ++v) sFk[v][kface_slot(rj, rk)] = F[v];
        }
      }
    }
    if (first) {                                   // uniform: no barrier skipped
      ring_load(I + 1);
      continue;
    }
    // the epilogue's state and volume loads go out before the barrier
    const unsigned c = (unsigned)I * sI + (unsigned)J * N[2] + K;
    const double* qi = qin.p[t.x];
    double vq[5], vz[5], ivol = 0.0;
    if (act) {
      ivol = ldu(B.vol, c);
#pragma unroll
      for (int v = 0; v < 5; ++v) vq[v] = ldu(qi + v * cells, c);
    }
    __syncthreads();
    if (act) {
      ivol = 1.0 / ivol;                           // (FP32-bulk form: one reciprocal)
      // outward sums in FP64 of the FP32 face values; the energy gets the gauge terms
      // e_ref x (mass divergence) + p_ref x (sum of w)
      double a[NFX];
      const int jup = tile_tid(tj + 1, tk), kup = kface_slot(tj, tk + 1);
#pragma unroll
      for (int v = 0; v < NFX; ++v) {
        double t = 0.0 - ((double)sFi[lo ^ 1][v][own_s] - (double)sFi[lo][v][own_s]);
        t = t - ((double)sFj[v][jup] - (double)sFj[v][tid]);
        t = t - ((double)sFk[v][kup] - (double)sFk[v][tid]);
        a[v] = t;
      }
      a[4] = (a[4] + B.e_ref * a[0]) + B.p_ref * a[5];
      double r[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) r[v] = a[v] * ivol;
      if (nd.n_sponge > 0) {
        const double xcv = ldu(XC, c);
        const int i = I - ng;
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
        const double* qz = q0.p[t.x];
#pragma unroll
        for (int v = 0; v < 5; ++v) vz[v] = ldu(qz + v * cells, c);
        const double a0 = MODE == 2 ? 0.75 : 1.0 / 3.0;
        const double a1 = MODE == 2 ? 0.25 : 2.0 / 3.0;
#pragma unroll
        for (int v = 0; v < 5; ++v) stu(o + v * cells, c, a0 * vz[v] + a1 * (vq[v] + dt * r[v]));
      }
    }
    // halo ring of the next plane, in flight while the next plane's staging waits.  No
    // barrier here: the next staging only overwrites what this plane's face tasks read
    // (all done before the barrier above), and the next face tasks write the exchange
    // buffers only after the next staging barrier, which every thread reaches after
    // its epilogue.
    if (I + 1 < ng + i1) ring_load(I + 1);
    lo ^= 1;
  }
}

}  // namespace wmx

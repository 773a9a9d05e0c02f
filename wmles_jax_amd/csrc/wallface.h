// Wall flux, tau_w and q_w of one WallModel face (wall_face), shared by the wall kernel and
// the side work of the gradient and primitive launches (reference bc.py:145-166).
#pragma once
#include "common.h"
#include "wallnet.h"

namespace wmx {

// Wall flux, tau_w and q_w of face idx0 of wall op `op` (bc.py:145-166); NET: the learned
// model with its weights staged in sW / sF, idx0 past the count clamped for the network.
template <bool NET>
__device__ __forceinline__ void wall_face(const BlockDev* __restrict__ blks, const WallOp& op,
                                          int idx0, const Ptrs& q, const NumDev& nd,
                                          const unsigned short* sW, const float* sF) {
  const int idx = NET ? min(idx0, op.count - 1) : idx0;
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
  if (NET) {
    double up;
    const bool fb = learned_wall(nd, sW, sF, 0, h, h / cbrt(B.vol[m]), U, wm.T, wm.p, f.T_wall,
                                 &tau, &qw, &up);
    const unsigned long long bal = __ballot(fb && idx0 < op.count);
    if (nd.net.nfb && (threadIdx.x & 63) == 0 && bal) atomicAdd(nd.net.nfb, (unsigned int)__popcll(bal));
    if (idx0 >= op.count) return;
  } else {
    equilibrium_wall(gas, nd.wm, h, U, wm.T, wm.p, f.T_wall, &tau, &qw);
  }
  double sg = s == 1 ? 1.0 : -1.0;
  double* W = B.wall[2 * a + s];
  const long long nw = (long long)B.n[t1] * B.n[t2];
  // FP32-bulk plans take the wall flux in gauge form: (p - p_ref) instead of p (the
  // p_ref S terms of a cell's faces telescope; no mass flux, so no energy gauge term)
  const double pw = nd.prec == WMX_PREC_FP32_BULK ? wb.p - B.p_ref : wb.p;
  W[idx] = sg * 0.0;
  for (int x = 0; x < 3; ++x) W[(1 + x) * nw + idx] = sg * ((pw * no[x] + tau * (up[x] / U)) * Sm);
  W[4 * nw + idx] = sg * ((-qw) * Sm);
  W[5 * nw + idx] = tau;
  W[6 * nw + idx] = qw;
}
// FP32-bulk gradient / SGS / Ducros.  One thread per run of grad_run(4) consecutive i
// cells of one (j, k) column: the i neighbours' phi and the shared i-face vectors stay in
// registers from one cell of the run to the next, so the planes i-1 / i+1 (far apart in
// memory) are read about once instead of three times.  Per cell the arithmetic is that
// of a single-cell evaluation.  Addressing: every plane / face-component array gets a
// wave-uniform base pointer and each load a 32-bit byte offset from it (the plan checks
// that every plane stays below 4 GiB), so loads take the scalar-base form and need one
// 32-bit offset register instead of a 64-bit address pair each.  The SGS model is a
// template parameter so that only one model is compiled into each instance.
// NaN-propagating min (numpy.min semantics)

}  // namespace wmx

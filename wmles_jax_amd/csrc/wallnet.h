// Learned equilibrium wall model: inference on one wavefront with BF16 MFMA.
//
// Network 3 -> 64 -> 64 -> 2, hidden activation ISRU f(a) = a / sqrt(1 + a^2) (one hardware
// reciprocal square root per unit).  Inputs (normalised): log Re_h, h/Delta,
// T_h/T_w.  Outputs (normalised): log u+ (u+ = U/u_tau) and log T+ (Kader's T+ at
// y+ = Re_h/u+), from which u_tau and q_w follow exactly as in the equilibrium model
// (reference wall_model.py:42-61).
//
// Layer 1 (3 inputs) runs in FP32 on the VALU; layers 2 and 3 run on
// v_mfma_f32_16x16x16bf16_1k with BF16 weights/activations and FP32 accumulation.
// Orientation hidden x points (D = W . H): the D tile a lane holds after layer l
// (rows 4*(lane/16)+v, column lane%16) is exactly the B operand of layer l+1's K chunk,
// so activations never leave registers.
#pragma once
#include "common.h"

namespace wmx {

constexpr int MLP_H = 64, MLP_IN = 3, MLP_OUT = 2;
// FP32 master parameters: W1 [64][3], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]
constexpr int MO_W1 = 0, MO_B1 = MO_W1 + MLP_H * MLP_IN, MO_W2 = MO_B1 + MLP_H,
              MO_B2 = MO_W2 + MLP_H * MLP_H, MO_W3 = MO_B2 + MLP_H, MO_B3 = MO_W3 + MLP_OUT * MLP_H,
              MLP_NPARAM = MO_B3 + MLP_OUT;
// BF16 inference copy: W2 [64][64] then W3 [2][64]
constexpr int MB_W2 = 0, MB_W3 = MB_W2 + MLP_H * MLP_H, MLP_NBF16 = MB_W3 + MLP_OUT * MLP_H;
// LDS image: BF16 rows padded to 72 entries (144 B) to spread the 8-byte A-operand reads
constexpr int MLP_LDW = 72;
constexpr int MLP_LDS_BF16 = (MLP_H + MLP_OUT) * MLP_LDW;
// FP32 inference copy: W1 [64][3] | b1 [64] | b2 [64] | b3 [2]
constexpr int MF_W1 = 0, MF_B1 = MF_W1 + MLP_H * MLP_IN, MF_B2 = MF_B1 + MLP_H,
              MF_B3 = MF_B2 + MLP_H, MLP_NF32 = MF_B3 + MLP_OUT;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// hidden activation of the network (inference): a / sqrt(1 + a^2) with v_rsq_f32
__device__ __forceinline__ float act_fast(float a) {
  return a * __builtin_amdgcn_rsqf(fmaf(a, a, 1.f));
}

__device__ __forceinline__ f32x4 mfma_bf16_16x16x16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Stage the network into LDS: sW [66][72] BF16, sF [322] FP32 (all threads of the
// block; the caller synchronises).
__device__ __forceinline__ void net_stage(const WallNetDev& net, unsigned short* sW, float* sF) {
  for (int i = threadIdx.x; i < MLP_NBF16; i += blockDim.x) {
    const int r = i / MLP_H, k = i % MLP_H;
    sW[r * MLP_LDW + k] = net.w[i];
  }
  for (int i = threadIdx.x; i < MLP_NF32; i += blockDim.x) sF[i] = net.f[i];
}

// Evaluate the network for the 64 points of this wavefront (one per lane).  Must be
// called by all 64 lanes.  x: this lane's normalised inputs; y: its normalised outputs.
__device__ __forceinline__ void net_eval(const unsigned short* __restrict__ sW,
                                         const float* __restrict__ sF, const float x[3],
                                         float y[2]) {
  const int l = threadIdx.x & 63, col = l & 15, grp = l >> 4;
  bf16x4 a2[4][4], a3[4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const unsigned short* p = sW + (MB_W2 / MLP_H + 16 * m + col) * MLP_LDW + 16 * c + 4 * grp;
#pragma unroll
      for (int v = 0; v < 4; ++v) a2[m][c][v] = (short)p[v];
    }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const unsigned short* p = sW + (MB_W3 / MLP_H + (col < MLP_OUT ? col : 0)) * MLP_LDW + 16 * c + 4 * grp;
#pragma unroll
    for (int v = 0; v < 4; ++v) a3[c][v] = col < MLP_OUT ? (short)p[v] : (short)0;
  }
  // this lane's 16 rows of layer 1 and the matching layer-2 biases stay in registers
  float w1[16][4], b2[16];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * c + 4 * grp + v;
      w1[4 * c + v][0] = sF[MF_W1 + row * MLP_IN];
      w1[4 * c + v][1] = sF[MF_W1 + row * MLP_IN + 1];
      w1[4 * c + v][2] = sF[MF_W1 + row * MLP_IN + 2];
      w1[4 * c + v][3] = sF[MF_B1 + row];
      b2[4 * c + v] = sF[MF_B2 + row];
    }
  float yo0 = 0.f, yo1 = 0.f;
#pragma unroll 1
  for (int r = 0; r < 4; ++r) {
    const int src = 16 * r + col;
    const float x0 = __shfl(x[0], src, 64), x1 = __shfl(x[1], src, 64), x2 = __shfl(x[2], src, 64);
    bf16x4 h1[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float* w = w1[4 * c + v];
        h1[c][v] = (short)f2bf(act_fast(w[3] + w[0] * x0 + w[1] * x1 + w[2] * x2));
      }
    bf16x4 h2[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 4; ++c) d = mfma_bf16_16x16x16(a2[m][c], h1[c], d);
#pragma unroll
      for (int v = 0; v < 4; ++v) h2[m][v] = (short)f2bf(act_fast(d[v] + b2[4 * m + v]));
    }
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; ++c) d = mfma_bf16_16x16x16(a3[c], h2[c], d);
    // rows 0 and 1 of D (the two outputs) sit in lanes 0..15, v = 0 and 1
    const float o0 = __shfl(d[0], col, 64), o1 = __shfl(d[1], col, 64);
    if (grp == r) {
      yo0 = o0 + sF[MF_B3];
      yo1 = o1 + sF[MF_B3 + 1];
    }
  }
  y[0] = yo0;
  y[1] = yo1;
}

// Spalding u+ (= U/u_tau) by the reference's safeguarded Newton from a given start
// (wall_model.py:19-40; the reference starts from clip(sqrt(Re_h), 1, 25)).  Stops at an
// exact fixed point of the update, which the remaining iterations would only repeat.
__device__ __forceinline__ double spalding_newton(const WallModelC& m, double Reh, double up, int iters) {
  const double k = m.kappa, ekb = m.ekb;
  for (int it = 0; it < iters; ++it) {
    const double ku = k * up;
    const double ex = exp(ku);
    const double poly2 = ((ex - 1.0) - ku) - 0.5 * (ku * ku);
    const double E = ekb * (poly2 - (ku * ku * ku) / 6.0);
    const double F = (up + E) - Reh / up;
    const double dF = (1.0 + (ekb * k) * poly2) + Reh / (up * up);
    const double nw = up - F / dF;
    const double upn = fmin(fmax(nw, 0.5 * up), 100.0);
    if (upn == up) break;
    up = upn;
  }
  return up;
}

__device__ __forceinline__ double spalding_residual(const WallModelC& m, double Reh, double up) {
  const double ku = m.kappa * up;
  const double poly2 = ((exp(ku) - 1.0) - ku) - 0.5 * (ku * ku);
  return (up + m.ekb * (poly2 - (ku * ku * ku) / 6.0)) - Reh / up;
}

// Kader T+ at y+ (wall_model.py:56-60)
__device__ __forceinline__ double kader_tplus(double Pr, double pr3, double beta, double yp) {
  yp = fmax(yp, 1e-6);
  const double py = Pr * yp, py2 = py * py;
  const double Gam = fmax((0.01 * (py2 * py2)) / (1.0 + (5.0 * pr3) * yp), 1e-20);
  return (Pr * yp) * exp(-Gam) + (2.12 * log(1.0 + yp) + beta) * exp(-1.0 / Gam);
}

// The learned model at one wall point, wave-cooperative (all 64 lanes call it).
//   mode 0: network start + n_polish Newton steps, checked by the Spalding residual;
//           q_w from Kader at the polished y+ (the step's form)
//   mode 1: network outputs only (u+ and T+ both from the network; diagnostics)
// Falls back to the full Newton of the reference when an input lies outside the
// training range or the relative residual |F(u+)|/u+ exceeds net.tol.  Returns
// (tau_w, q_w, u+) and whether the point fell back.
__device__ __forceinline__ bool learned_wall(const NumDev& nd, const unsigned short* sW,
                                             const float* sF, int mode, double h, double hdel,
                                             double U, double T_h, double p_h, double Tw_in,
                                             double* tau_w, double* q_w, double* uplus) {
  const GasC& g = nd.gas;
  const WallNetDev& net = nd.net;
  U = fmax(U, 1e-8);
  const bool adiabatic = Tw_in < 0.0;
  const double T_w = adiabatic ? T_h + (nd.wm.pr13 * (U * U)) / (2.0 * g.cp) : Tw_in + 0.0 * T_h;
  const double rho_w = p_h / (g.R * T_w);
  const double mu_w = mu_suth(g, T_w);
  const double Reh = ((h * U) * rho_w) / mu_w;
  const double lr = log(Reh), tr = T_h / T_w;
  // the equilibrium law does not depend on h/Delta: clamp it into the trained range
  const double hd = fmin(fmax(hdel, net.lo[1]), net.hi[1]);
  const float x[3] = {(float)((lr - net.xc[0]) / net.xs[0]), (float)((hd - net.xc[1]) / net.xs[1]),
                      (float)((tr - net.xc[2]) / net.xs[2])};
  float y[2];
  net_eval(sW, sF, x, y);
  double up = exp((double)y[0] * net.ys[0] + net.yc[0]);
  bool fb = !(lr >= net.lo[0] && lr <= net.hi[0] && tr >= net.lo[2] && tr <= net.hi[2]) ||
            !(up > 0.5 && up < 100.0);
  double Tp = -1.0;
  if (mode == 0 && !fb) {
    up = spalding_newton(nd.wm, Reh, up, net.n_polish);
    fb = !(fabs(spalding_residual(nd.wm, Reh, up)) <= net.tol * up);
  } else if (mode == 1 && !fb) {
    Tp = exp((double)y[1] * net.ys[1] + net.yc[1]);
  }
  if (fb) {
    up = spalding_newton(nd.wm, Reh, fmin(fmax(sqrt(Reh), 1.0), 25.0), nd.wm.n_iter);
    Tp = -1.0;
  }
  const double ut = U / up;
  const double tau = rho_w * (ut * ut);
  *tau_w = tau;
  *uplus = up;
  if (adiabatic) {
    *q_w = 0.0 * tau;
  } else {
    const double yp = fmax(((h * ut) * rho_w) / mu_w, 1e-6);
    if (Tp < 0.0) Tp = kader_tplus(g.Pr, nd.wm.pr3, nd.wm.beta, yp);
    *q_w = (((rho_w * g.cp) * ut) * (T_w - T_h)) / Tp;
  }
  return fb;
}

}  // namespace wmx

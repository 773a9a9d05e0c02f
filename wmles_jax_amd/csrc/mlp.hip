// Learned wall model: training data, FP32 training kernels and the point-array
// evaluation of the BF16 MFMA network (the network itself is in wallnet.h).
//
//   k_wm_dataset      samples of the method itself: Spalding u+ by the reference's
//                     safeguarded 30-step Newton, Kader T+ (wall_model.py:19-61)
//   k_mlp_train       FP32 forward + backward of 64 samples per workgroup; per-group
//                     gradient partials (deterministic, no atomics)
//   k_mlp_reduce_adam sum of the partials, optional Adam update of the FP32 master weights
//   k_mlp_pack        FP32 master -> BF16 / FP32 inference copies
//   k_mlp_forward32   FP32 forward (the reference the BF16 path is compared with)
//   k_wm_net_eval     the learned model on point arrays (wall-model callable boundary)
//   k_mfma_probe      one 16x16x16 BF16 MFMA tile with the lane layout wallnet.h assumes
#include "common.h"
#include "wallnet.h"

namespace wmx {

// counter-based uniform in [0, 1): splitmix64 finaliser of (seed, 4 i + stream + 1)
__device__ __forceinline__ double uhash(unsigned long long seed, long long i, int s) {
  unsigned long long x = seed ^ (0x9e3779b97f4a7c15ULL * (unsigned long long)(4 * i + s + 1));
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// -------------------------------------------------------------------------- dataset
// Features uniform in the trained box: log Re_h, h/Delta, T_h/T_w.  Targets log u+ and
// log T+ with T+ = Kader(y+ = Re_h / u+).  X, Y normalised FP32; raw (optional) FP64
// rows (log Re_h, h/Delta, T_h/T_w, u+, T+).
__global__ void k_wm_dataset(NumDev nd, long long n, unsigned long long seed,
                             float* __restrict__ X, float* __restrict__ Y, double* __restrict__ raw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const WallNetDev& net = nd.net;
  double f[3];
  for (int k = 0; k < 3; ++k) f[k] = net.lo[k] + (net.hi[k] - net.lo[k]) * uhash(seed, i, k);
  const double Reh = exp(f[0]);
  const double up = spalding_newton(nd.wm, Reh, fmin(fmax(sqrt(Reh), 1.0), 25.0), nd.wm.n_iter);
  const double Tp = kader_tplus(nd.gas.Pr, nd.wm.pr3, nd.wm.beta, Reh / up);
  for (int k = 0; k < 3; ++k) X[3 * i + k] = (float)((f[k] - net.xc[k]) / net.xs[k]);
  Y[2 * i + 0] = (float)((log(up) - net.yc[0]) / net.ys[0]);
  Y[2 * i + 1] = (float)((log(Tp) - net.yc[1]) / net.ys[1]);
  if (raw) {
    for (int k = 0; k < 3; ++k) raw[5 * i + k] = f[k];
    raw[5 * i + 3] = up;
    raw[5 * i + 4] = Tp;
  }
}

// hidden activation (training and the FP32 forward): a / sqrt(1 + a^2); its derivative in
// terms of the activation h is (1 - h^2)^(3/2)
__device__ __forceinline__ float act32(float a) { return a / sqrtf(1.f + a * a); }
__device__ __forceinline__ float act32_prime(float h) {
  const float u = 1.f - h * h;
  return u * sqrtf(u);
}

// ---------------------------------------------------------------------------- train
// 64 samples per workgroup of 4 waves.  Lane = sample, wave w owns hidden units
// w, w+4, ..., so weight reads are LDS broadcasts and activation reads walk rows of
// stride 65.  Loss = mean squared error over the batch x 2 outputs.
constexpr int TG = 64, TLD = MLP_H + 1;

__global__ __launch_bounds__(256) void k_mlp_train(const float* __restrict__ P,
                                                   const float* __restrict__ X,
                                                   const float* __restrict__ Y, long long off,
                                                   int batch, float* __restrict__ gpart,
                                                   float* __restrict__ lpart) {
  __shared__ float sP[MLP_NPARAM];
  __shared__ float sH1[TG * TLD], sH2[TG * TLD], sD1[TG * TLD], sD2[TG * TLD];
  __shared__ float sX[TG * 3], sY[TG * 2], sDY[TG * 2];
  __shared__ float sL[4];
  const int t = threadIdx.x, q = t & 63, w = t >> 6;
  const int s0 = blockIdx.x * TG;
  const int nvalid = min(TG, batch - s0);
  for (int i = t; i < MLP_NPARAM; i += 256) sP[i] = P[i];
  if (t < TG * 3) sX[t] = (t / 3) < nvalid ? X[3 * (off + s0) + t] : 0.f;
  if (t < TG * 2) {
    sY[t] = (t / 2) < nvalid ? Y[2 * (off + s0) + t] : 0.f;
    sDY[t] = 0.f;
  }
  __syncthreads();
  const bool ok = q < nvalid;
  // layer 1
  for (int j = w; j < MLP_H; j += 4) {
    float a = sP[MO_B1 + j];
    for (int i = 0; i < MLP_IN; ++i) a += sP[MO_W1 + j * MLP_IN + i] * sX[q * 3 + i];
    sH1[q * TLD + j] = act32(a);
  }
  __syncthreads();
  // layer 2
  for (int j = w; j < MLP_H; j += 4) {
    float a = sP[MO_B2 + j];
#pragma unroll 16
    for (int k = 0; k < MLP_H; ++k) a += sP[MO_W2 + j * MLP_H + k] * sH1[q * TLD + k];
    sH2[q * TLD + j] = act32(a);
  }
  __syncthreads();
  // output layer, d loss / d y
  float err2 = 0.f;
  if (w < MLP_OUT) {
    const int o = w;
    float a = sP[MO_B3 + o];
#pragma unroll 16
    for (int k = 0; k < MLP_H; ++k) a += sP[MO_W3 + o * MLP_H + k] * sH2[q * TLD + k];
    const float e = ok ? a - sY[q * 2 + o] : 0.f;
    err2 = e * e;
    sDY[q * 2 + o] = 2.f * e / (float)(batch * MLP_OUT);
  }
  __syncthreads();
  // back through layer 2's activation
  for (int k = w; k < MLP_H; k += 4) {
    float g = 0.f;
    for (int o = 0; o < MLP_OUT; ++o) g += sP[MO_W3 + o * MLP_H + k] * sDY[q * 2 + o];
    sD2[q * TLD + k] = g * act32_prime(sH2[q * TLD + k]);
  }
  __syncthreads();
  // back through layer 1's activation
  for (int k = w; k < MLP_H; k += 4) {
    float g = 0.f;
#pragma unroll 16
    for (int j = 0; j < MLP_H; ++j) g += sP[MO_W2 + j * MLP_H + k] * sD2[q * TLD + j];
    sD1[q * TLD + k] = g * act32_prime(sH1[q * TLD + k]);
  }
  __syncthreads();
  // weight gradients summed over the group's samples (invalid samples carry zero deltas)
  float* gp = gpart + (long long)blockIdx.x * MLP_NPARAM;
  for (int e = t; e < MLP_H * MLP_H; e += 256) {
    const int j = e / MLP_H, k = e % MLP_H;
    float a = 0.f;
#pragma unroll 16
    for (int s = 0; s < TG; ++s) a += sD2[s * TLD + j] * sH1[s * TLD + k];
    gp[MO_W2 + e] = a;
  }
  constexpr int NSMALL = MLP_H * MLP_IN + 2 * MLP_H + MLP_OUT * MLP_H + MLP_OUT;
  for (int e = t; e < NSMALL; e += 256) {
    float a = 0.f;
    int dst;
    if (e < MLP_H * MLP_IN) {                                        // W1
      const int j = e / MLP_IN, i = e % MLP_IN;
      for (int s = 0; s < TG; ++s) a += sD1[s * TLD + j] * sX[s * 3 + i];
      dst = MO_W1 + e;
    } else if (e < MLP_H * MLP_IN + MLP_H) {                         // b1
      const int j = e - MLP_H * MLP_IN;
      for (int s = 0; s < TG; ++s) a += sD1[s * TLD + j];
      dst = MO_B1 + j;
    } else if (e < MLP_H * MLP_IN + 2 * MLP_H) {                     // b2
      const int j = e - MLP_H * MLP_IN - MLP_H;
      for (int s = 0; s < TG; ++s) a += sD2[s * TLD + j];
      dst = MO_B2 + j;
    } else if (e < MLP_H * MLP_IN + 2 * MLP_H + MLP_OUT * MLP_H) {   // W3
      const int l = e - MLP_H * MLP_IN - 2 * MLP_H, o = l / MLP_H, k = l % MLP_H;
      for (int s = 0; s < TG; ++s) a += sDY[s * 2 + o] * sH2[s * TLD + k];
      dst = MO_W3 + l;
    } else {                                                         // b3
      const int o = e - MLP_H * MLP_IN - 2 * MLP_H - MLP_OUT * MLP_H;
      for (int s = 0; s < TG; ++s) a += sDY[s * 2 + o];
      dst = MO_B3 + o;
    }
    gp[dst] = a;
  }
  // loss partial (sum of squared errors of the group)
  for (int o = 32; o > 0; o >>= 1) err2 += __shfl_xor(err2, o, 64);
  if (q == 0) sL[w] = err2;
  __syncthreads();
  if (t == 0) lpart[blockIdx.x] = ((sL[0] + sL[1]) + sL[2]) + sL[3];
}

__global__ void k_mlp_reduce_adam(float* __restrict__ P, const float* __restrict__ gpart,
                                  int ngroups, float* __restrict__ gout, float* __restrict__ m,
                                  float* __restrict__ v, int step, float lr, int update) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  float g = 0.f;
  for (int q = 0; q < ngroups; ++q) g += gpart[(long long)q * MLP_NPARAM + i];
  if (gout) gout[i] = g;
  if (update) {   // Adam (Kingma & Ba 2015): beta1 0.9, beta2 0.999, eps 1e-8
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)step));
    const float vh = vi / (1.f - powf(b2, (float)step));
    P[i] = P[i] - lr * mh / (sqrtf(vh) + eps);
  }
}

__global__ void k_mlp_pack(const float* __restrict__ P, unsigned short* __restrict__ W,
                           float* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < MLP_H * MLP_H) W[MB_W2 + i] = f2bf(P[MO_W2 + i]);
  if (i < MLP_OUT * MLP_H) W[MB_W3 + i] = f2bf(P[MO_W3 + i]);
  if (i < MLP_H * MLP_IN) F[MF_W1 + i] = P[MO_W1 + i];
  if (i < MLP_H) {
    F[MF_B1 + i] = P[MO_B1 + i];
    F[MF_B2 + i] = P[MO_B2 + i];
  }
  if (i < MLP_OUT) F[MF_B3 + i] = P[MO_B3 + i];
}

__global__ __launch_bounds__(256) void k_mlp_forward32(const float* __restrict__ P,
                                                       const float* __restrict__ X, long long n,
                                                       float* __restrict__ Yo) {
  __shared__ float sP[MLP_NPARAM];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) sP[i] = P[i];
  __syncthreads();
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float x0 = X[3 * s], x1 = X[3 * s + 1], x2 = X[3 * s + 2];
  float h1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; ++j)
    h1[j] = act32(sP[MO_B1 + j] + sP[MO_W1 + 3 * j] * x0 + sP[MO_W1 + 3 * j + 1] * x1 +
                  sP[MO_W1 + 3 * j + 2] * x2);
  float y0 = sP[MO_B3], y1 = sP[MO_B3 + 1];
  for (int j = 0; j < MLP_H; ++j) {
    float a = sP[MO_B2 + j];
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) a += sP[MO_W2 + j * MLP_H + k] * h1[k];
    const float h = act32(a);
    y0 += sP[MO_W3 + j] * h;
    y1 += sP[MO_W3 + MLP_H + j] * h;
  }
  Yo[2 * s] = y0;
  Yo[2 * s + 1] = y1;
}

// The learned model on point arrays; T_wall < 0: adiabatic.  hdel may be null (the
// middle of the trained h/Delta range is used; the equilibrium law ignores it).
__global__ __launch_bounds__(256) void k_wm_net_eval(NumDev nd, int mode, long long n,
                                                     const double* __restrict__ h,
                                                     const double* __restrict__ hdel,
                                                     const double* __restrict__ U,
                                                     const double* __restrict__ Th,
                                                     const double* __restrict__ ph, double T_wall,
                                                     double* __restrict__ tau,
                                                     double* __restrict__ qw,
                                                     double* __restrict__ uplus) {
  __shared__ unsigned short sW[MLP_LDS_BF16];
  __shared__ float sF[MLP_NF32];
  net_stage(nd.net, sW, sF);
  __syncthreads();
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if ((long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63) >= n) return;   // whole waves
  const long long i = i0 < n ? i0 : n - 1;
  const double hd = hdel ? hdel[i] : 0.5 * (nd.net.lo[1] + nd.net.hi[1]);
  double t, q, up;
  const bool fb = learned_wall(nd, sW, sF, mode, h[i], hd, U[i], Th[i], ph[i], T_wall, &t, &q, &up);
  const unsigned long long bal = __ballot(fb && i0 < n);
  if (nd.net.nfb && (threadIdx.x & 63) == 0 && bal) atomicAdd(nd.net.nfb, (unsigned int)__popcll(bal));
  if (i0 >= n) return;
  tau[i] = t;
  qw[i] = q;
  if (uplus) uplus[i] = up;
}

__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B,
                             float* __restrict__ D) {
  const int l = threadIdx.x, col = l & 15, grp = l >> 4;
  bf16x4 a, b;
  for (int v = 0; v < 4; ++v) {
    a[v] = (short)f2bf(A[col * 16 + 4 * grp + v]);       // A[i = col][k = 4 grp + v]
    b[v] = (short)f2bf(B[(4 * grp + v) * 16 + col]);     // B[k = 4 grp + v][j = col]
  }
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  d = mfma_bf16_16x16x16(a, b, d);
  for (int v = 0; v < 4; ++v) D[(4 * grp + v) * 16 + col] = d[v];   // D[i = 4 grp + v][j = col]
}

// ------------------------------------------------------------------------ launchers
hipError_t launch_wm_dataset(const NumDev& nd, long long n, unsigned long long seed, float* X,
                             float* Y, double* raw, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wm_dataset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nd, n, seed,
                     X, Y, raw);
  return hipGetLastError();
}

hipError_t launch_mlp_train(float* P, const float* X, const float* Y, long long off, int batch,
                            float* gpart, float* lpart, float* gout, float* m, float* v, int step,
                            float lr, int update, hipStream_t s) {
  const int ng = (batch + TG - 1) / TG;
  hipLaunchKernelGGL(k_mlp_train, dim3(ng), dim3(256), 0, s, P, X, Y, off, batch, gpart, lpart);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, P, gpart,
                     ng, gout, m, v, step, lr, update);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* P, unsigned short* W, float* F, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, s, P, W, F);
  return hipGetLastError();
}

hipError_t launch_mlp_forward32(const float* P, const float* X, long long n, float* Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, X, n, Y);
  return hipGetLastError();
}

hipError_t launch_wm_net_eval(const NumDev& nd, int mode, long long n, const double* h,
                              const double* hdel, const double* U, const double* Th,
                              const double* ph, double T_wall, double* tau, double* qw,
                              double* uplus, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_wm_net_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nd, mode, n,
                     h, hdel, U, Th, ph, T_wall, tau, qw, uplus);
  return hipGetLastError();
}

hipError_t launch_mfma_probe(const float* A, const float* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

int mlp_nparam() { return MLP_NPARAM; }
int mlp_nbf16() { return MLP_NBF16; }
int mlp_nf32() { return MLP_NF32; }

}  // namespace wmx

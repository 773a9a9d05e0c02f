// Ghost-cell / ghost-node fills and block geometry.
//
// Built with -ffp-contract=off: every value here is a copy or a short expression
// evaluated in the reference's association order, so the ghost maps and the metric
// terms agree with the reference bit for bit.
#include "ghostfill.h"

namespace wmx {

__global__ void k_fill_nodes(const BlockDev* __restrict__ blks, const GhostOp* __restrict__ ops,
                             Ptrs X) {
  const GhostOp op = ops[blockIdx.y];
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= op.count) return;
  const BlockDev& B = blks[op.blk];
  const int ng = B.ng;
  int dims[3] = {B.N[0] + 1, B.N[1] + 1, B.N[2] + 1};
  int m, c[3];
  ghost_coords(idx, op.axis, ng, dims, &m, c);
  const int a = op.axis, s = op.side, n = B.n[a];
  c[a] = s == 1 ? ng + n + 1 + m : ng - 1 - m;
  double* Xd = X.p[op.blk];
  const long long plane = (long long)dims[0] * dims[1] * dims[2];
  const long long di = ((long long)c[0] * dims[1] + c[1]) * dims[2] + c[2];
  const wmx_face_desc& f = B.faces[2 * a + s];
  if (f.kind == WMX_FACE_INTERFACE) {
    const BlockDev& S = blks[f.nb];
    int sd[3] = {S.N[0] + 1, S.N[1] + 1, S.N[2] + 1};
    int cs[3];
    interface_src(f, a, s, m, c, dims, ng, S.n[f.src_axis], 1, cs);
    const long long si = ((long long)cs[0] * sd[1] + cs[1]) * sd[2] + cs[2];
    const long long splane = (long long)sd[0] * sd[1] * sd[2];
    const double* Xs = X.p[f.nb];
    for (int v = 0; v < 3; ++v) Xd[v * plane + di] = Xs[v * splane + si] + f.shift[v];
  } else {
    // point reflection through the boundary node plane, reference mesh.py:158-166
    int cb[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
    if (s == 0) { cb[a] = ng; cm[a] = ng + 1 + m; }
    else { cb[a] = ng + n; cm[a] = ng + n - 1 - m; }
    const long long bi = ((long long)cb[0] * dims[1] + cb[1]) * dims[2] + cb[2];
    const long long mi = ((long long)cm[0] * dims[1] + cm[1]) * dims[2] + cm[2];
    for (int v = 0; v < 3; ++v) Xd[v * plane + di] = 2.0 * Xd[v * plane + bi] - Xd[v * plane + mi];
  }
}

// Ghost-cell fill of one phase of the sequential sweep, reference solver.py:59-73,
// mesh.py:125-155 (interfaces) and bc.py:14-143 (mirror fills); the per-cell value is
// ghost_value (ghostfill.h).
__global__ void k_fill_cells(const BlockDev* __restrict__ blks, const GhostOp* __restrict__ ops,
                             Ptrs q, GasC gas) {
  fill_cell(blks, ops[blockIdx.y], q, gas, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// Face vectors and face centres of direction d, reference mesh.py:58-65.
__global__ void k_geom_faces(const double* __restrict__ X, int N0, int N1, int N2, int d,
                             double* __restrict__ S, double* __restrict__ xf) {
  const int N[3] = {N0, N1, N2};
  const int M0 = N0 + (d == 0), M1 = N1 + (d == 1), M2 = N2 + (d == 2);
  long long tot = (long long)M0 * M1 * M2;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= tot) return;
  int K = (int)(idx % M2);
  long long r = idx / M2;
  int J = (int)(r % M1);
  int I = (int)(r / M1);
  const int X1 = N1 + 1, X2 = N2 + 1;
  const long long xp = (long long)(N0 + 1) * X1 * X2;
  // corner(a, b): d=0 -> X[:, I, J+a, K+b]; d=1 -> X[:, I+b, J, K+a]; d=2 -> X[:, I+a, J+b, K]
  auto corner = [&](int a, int b, double* v) {
    int i = I, j = J, k = K;
    if (d == 0) { j += a; k += b; }
    else if (d == 1) { i += b; k += a; }
    else { i += a; j += b; }
    long long o = ((long long)i * X1 + j) * X2 + k;
    v[0] = X[o]; v[1] = X[xp + o]; v[2] = X[2 * xp + o];
  };
  double c00[3], c10[3], c01[3], c11[3];
  corner(0, 0, c00); corner(1, 0, c10); corner(0, 1, c01); corner(1, 1, c11);
  double a0 = c11[0] - c00[0], a1 = c11[1] - c00[1], a2 = c11[2] - c00[2];
  double b0 = c01[0] - c10[0], b1 = c01[1] - c10[1], b2 = c01[2] - c10[2];
  double cr0 = a1 * b2 - a2 * b1, cr1 = a2 * b0 - a0 * b2, cr2 = a0 * b1 - a1 * b0;
  S[idx] = 0.5 * cr0; S[tot + idx] = 0.5 * cr1; S[2 * tot + idx] = 0.5 * cr2;
  for (int v = 0; v < 3; ++v) xf[v * tot + idx] = 0.25 * (((c00[v] + c10[v]) + c01[v]) + c11[v]);
  (void)N;
}

// Divergence-theorem volume and 8-node centroid, reference mesh.py:67-73.
__global__ void k_geom_cells(const double* __restrict__ X, int N0, int N1, int N2,
                             const double* __restrict__ S0, const double* __restrict__ S1,
                             const double* __restrict__ S2, const double* __restrict__ x0,
                             const double* __restrict__ x1, const double* __restrict__ x2,
                             double* __restrict__ vol, double* __restrict__ xc) {
  const int N[3] = {N0, N1, N2};
  long long tot = (long long)N0 * N1 * N2;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= tot) return;
  int K = (int)(idx % N2);
  long long r = idx / N2;
  int J = (int)(r % N1);
  int I = (int)(r / N1);
  auto dotf = [&](const double* S, const double* x, int d, int i, int j, int k) {
    long long fc = fcount(N, d);
    long long o = fidx(N, d, i, j, k);
    return (x[o] * S[o] + x[fc + o] * S[fc + o]) + x[2 * fc + o] * S[2 * fc + o];
  };
  double v = (((((dotf(S0, x0, 0, I + 1, J, K) - dotf(S0, x0, 0, I, J, K)) +
                 dotf(S1, x1, 1, I, J + 1, K)) - dotf(S1, x1, 1, I, J, K)) +
               dotf(S2, x2, 2, I, J, K + 1)) - dotf(S2, x2, 2, I, J, K)) / 3.0;
  vol[idx] = v;
  const int X1 = N1 + 1, X2 = N2 + 1;
  const long long xp = (long long)(N0 + 1) * X1 * X2;
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int e = 0; e < 2; ++e)
          acc = acc + X[c * xp + ((long long)(I + a) * X1 + (J + b)) * X2 + (K + e)];
    xc[c * tot + idx] = acc / 8.0;
  }
}

// numpy.interp semantics for an increasing table column (left/right clamp), split in two:
// where x falls in the abscissa column (the same for every column of a table, so one
// search serves them all), then the value of one column there.
struct InterpAt {
  int lo;
  bool exact;                                  // fp[lo] itself (clamped, or x on a node)
};

__device__ InterpAt interp_at(double x, const double* xp, int stride, int n) {
  if (x < xp[0]) return {0, true};
  if (x > xp[(n - 1) * stride]) return {n - 1, true};
  if (x == xp[(n - 1) * stride]) return {n - 1, true};
  int lo = 0, hi = n - 1;                      // xp[lo] <= x < xp[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (xp[mid * stride] <= x) lo = mid; else hi = mid;
  }
  return {lo, xp[lo * stride] == x || n < 2};   // n < 2: only a NaN x gets here
}

__device__ double interp_col(InterpAt at, double x, const double* xp, const double* fp, int stride) {
  const int lo = at.lo;
  if (at.exact) return fp[lo * stride];
  double slope = (fp[(lo + 1) * stride] - fp[lo * stride]) / (xp[(lo + 1) * stride] - xp[lo * stride]);
  return slope * (x - xp[lo * stride]) + fp[lo * stride];
}

// Synthetic inflow ghost fill, reference smooth_ramp.py:150-197.  One workgroup per
// (ghost layer m, row J); threads over K (full ghosted extent, <= blockDim).
__global__ void k_inflow(const BlockDev* __restrict__ blks, int blk, InflowDev in, Ptrs q,
                         GasC gas, double time, const double* __restrict__ d_dt, double t_coef) {
  extern __shared__ double red[];
  const BlockDev& B = blks[blk];
  const int ng = B.ng;
  const int m = blockIdx.y;
  const int J = blockIdx.x;
  const int K = threadIdx.x;
  const int N2 = B.N[2], nk = B.n[2];
  const int gi = ng - 1 - m;
  const bool valid = K < N2;
  double t = time;                              // host stage time, or base time + c dt
  if (d_dt) t = time + t_coef * d_dt[B.case_id];
  const long long ci = cidx(B.N, gi, J, valid ? K : 0);
  const long long cells = B.cells;
  double y_l = B.xc[cells + ci] / in.L;
  double z_l = B.xc[2 * cells + ci] / in.L;
  double X0 = in.U_inf * t / in.L;
  const bool inner = valid && K >= ng && K < ng + nk;
  // the profile columns at y_l, loaded before the mode sums (one search of the y column)
  const double* tb = in.table;
  const InterpAt at = interp_at(y_l, tb, 7, in.n_tab);
  double tcol[6];
  for (int c = 0; c < 6; ++c) tcol[c] = interp_col(at, y_l, tb, tb + 1 + c, 7);
  const int nm = in.n_modes;
  // cos of every (cell, mode) phase, computed by all threads of the workgroup into LDS
  // (when the row fits), then summed per cell in mode order: the same values and the same
  // sum as one thread per cell doing its modes in turn.  Mode-major with an odd pitch, so
  // that lanes of one mode (the sums) or of one cell (the phases) fall in distinct banks.
  const int P2 = N2 | 1;
  double* cosb = red + blockDim.x + 1;                  // [nm][P2] when cooperative
  double* yz = cosb + (long long)P2 * nm;               // [2][N2]
  const bool coop = in.coop;
  if (coop) {
    for (int kk = threadIdx.x; kk < N2; kk += blockDim.x) {
      const long long cj = cidx(B.N, gi, J, kk);
      yz[kk] = B.xc[cells + cj] / in.L;
      yz[N2 + kk] = B.xc[2 * cells + cj] / in.L;
    }
  }
  double fld[3];
  for (int comp = 0; comp < 3; ++comp) {
    const double* md = in.modes + (long long)comp * 4 * nm;
    double f = 0.0;
    if (coop) {
      __syncthreads();                                   // yz staged / previous sums done
      if ((int)blockDim.x % nm == 0) {
        // every element of this thread has the same mode k: its coefficients (and the
        // time term) are loaded once; the phase keeps its operation order
        const int k = (int)threadIdx.x % nm;
        const double ax = md[k] * X0, by = md[nm + k], cz = md[2 * nm + k], dk = md[3 * nm + k];
        for (int kk = (int)threadIdx.x / nm; kk < N2; kk += (int)blockDim.x / nm) {
          const double ph = ((ax + by * yz[kk]) + cz * yz[N2 + kk]) + dk;
          cosb[k * P2 + kk] = cos(ph);
        }
      } else {
        for (int e = threadIdx.x; e < N2 * nm; e += blockDim.x) {
          const int kk = e / nm, k = e - kk * nm;
          const double ph = ((md[k] * X0 + md[nm + k] * yz[kk]) + md[2 * nm + k] * yz[N2 + kk]) + md[3 * nm + k];
          cosb[k * P2 + kk] = cos(ph);
        }
      }
      __syncthreads();
      if (valid) {
        double acc = cosb[K];
        for (int k = 1; k < nm; ++k) acc = acc + cosb[k * P2 + K];
        f = acc;
      }
    } else if (valid) {
      double acc = 0.0;
      for (int k = 0; k < nm; ++k) {
        double ph = ((md[k] * X0 + md[nm + k] * y_l) + md[2 * nm + k] * z_l) + md[3 * nm + k];
        acc = (k == 0) ? cos(ph) : acc + cos(ph);
      }
      f = acc;
    }
    fld[comp] = f;
  }
  if (coop) __syncthreads();                             // the row statistics reuse red[]
  // row statistics over the interior span, summed in numpy's add.reduce order
  auto row_mean = [&](double val) {
    red[K] = inner ? val : 0.0;
    __syncthreads();
    if (K == 0) red[blockDim.x] = np_sum([&](int i) { return red[ng + i]; }, nk) / nk;
    __syncthreads();
    double r = red[blockDim.x];
    __syncthreads();
    return r;
  };
  double mean0 = row_mean(fld[0]);
  fld[0] = fld[0] - mean0;
  double var0 = row_mean(fld[0] * fld[0]);
  double n0 = fld[0] / sqrt(fmax(var0, 1e-20));
  double mean1 = row_mean(fld[1]);
  fld[1] = fld[1] - mean1;
  double cov = row_mean(fld[1] * n0);
  fld[1] = fld[1] - cov * n0;
  double var1 = row_mean(fld[1] * fld[1]);
  double n1 = fld[1] / sqrt(fmax(var1, 1e-20));
  double mean2 = row_mean(fld[2]);
  fld[2] = fld[2] - mean2;
  double var2 = row_mean(fld[2] * fld[2]);
  double n2 = fld[2] / sqrt(fmax(var2, 1e-20));
  if (!valid) return;
  auto col = [&](int c) { return tcol[c - 1]; };
  double Ur = in.U_inf, U2 = Ur * Ur;
  double uu = fmax(col(3), 0.0) * U2, vv = fmax(col(4), 0.0) * U2, ww = fmax(col(5), 0.0) * U2;
  double uv = col(6) * U2;
  double su = sqrt(uu);
  double l10 = su > 1e-14 ? uv / fmax(su, 1e-14) : 0.0;
  double l11 = sqrt(fmax(vv - l10 * l10, 0.0));
  double u = Ur * col(1) + su * n0;
  double v = (Ur * col(2) + l10 * n0) + l11 * n1;
  double w = sqrt(ww) * n2;
  double out[5];
  prim_to_cons(gas, in.rho_inf + 0.0, u, v, w, in.p_inf + 0.0, out);
  double* qd = q.p[blk];
  for (int c = 0; c < 5; ++c) qd[c * cells + ci] = out[c];
}

// ----------------------------------------------------------------------- launchers
hipError_t launch_fill_nodes(const BlockDev* d_blks, const GhostOp* d_ops, int nops, int maxcount,
                             Ptrs X, hipStream_t s) {
  if (nops <= 0 || maxcount <= 0) return hipSuccess;
  dim3 grid((maxcount + 255) / 256, nops);
  hipLaunchKernelGGL(k_fill_nodes, grid, dim3(256), 0, s, d_blks, d_ops, X);
  return hipGetLastError();
}

hipError_t launch_fill_cells(const BlockDev* d_blks, const GhostOp* d_ops, int nops, int maxcount,
                             Ptrs q, GasC gas, hipStream_t s) {
  if (nops <= 0 || maxcount <= 0) return hipSuccess;
  dim3 grid((maxcount + 255) / 256, nops);
  hipLaunchKernelGGL(k_fill_cells, grid, dim3(256), 0, s, d_blks, d_ops, q, gas);
  return hipGetLastError();
}

hipError_t launch_inflow(const BlockDev* d_blks, int blk, int N1, int N2, int ng, InflowDev in,
                         Ptrs q, GasC gas, double time, const double* d_dt, double coef,
                         hipStream_t s) {
  int threads = ((N2 + 63) / 64) * 64;
  if (threads > 1024) return hipErrorInvalidValue;
  dim3 grid(N1, ng);
  // cooperative phase evaluation when a row's (cell, mode) table fits in 56 KB of LDS
  const size_t coop_bytes = sizeof(double) * ((size_t)(N2 | 1) * in.n_modes + 2 * (size_t)N2);
  in.coop = coop_bytes <= 56 * 1024 ? 1 : 0;
  if (in.coop && threads < 256) threads = 256;           // more lanes for the phase table
  size_t shm = sizeof(double) * (threads + 1) + (in.coop ? coop_bytes : 0);
  hipLaunchKernelGGL(k_inflow, grid, dim3(threads), shm, s, d_blks, blk, in, q, gas, time, d_dt,
                     coef);
  return hipGetLastError();
}

hipError_t launch_geometry(const double* X, const int* N, double* vol, double* xc, double* const* S,
                           double* const* xf, hipStream_t s) {
  for (int d = 0; d < 3; ++d) {
    long long tot = (long long)(N[0] + (d == 0)) * (N[1] + (d == 1)) * (N[2] + (d == 2));
    hipLaunchKernelGGL(k_geom_faces, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, X, N[0],
                       N[1], N[2], d, S[d], xf[d]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  long long tot = (long long)N[0] * N[1] * N[2];
  hipLaunchKernelGGL(k_geom_cells, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, X, N[0],
                     N[1], N[2], S[0], S[1], S[2], xf[0], xf[1], xf[2], vol, xc);
  return hipGetLastError();
}

}  // namespace wmx

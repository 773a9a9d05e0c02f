// Internal device-side descriptors shared by the kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wmles_amd.h"
#include "physics.h"

namespace wmx {

constexpr int NP = 8;    // prim planes: rho u v w p T mu e_int
constexpr int NGR = 15;  // gradient record planes: g[4][3], mu_eff, k_eff, theta
constexpr int WALL_PLANES = 7;  // Fw[5], tau, qw

enum PrimPlane { P_R = 0, P_U, P_V, P_W, P_P, P_T, P_MU, P_E };
enum GradPlane { G_MU = 12, G_K = 13, G_TH = 14 };
// FP64 records are plane-major (plane v of cell c at v * cells + c).  FP32-bulk records
// move in 16-byte groups: the primitive record is two planes of float4,
// Pf[4 c ..] = (u, v, w, T') and Pf[4 cells + 4 c ..] = (rho, p', mu, e'), so the
// gradient kernel reads one group per neighbour; the gradient record is four planes of
// float4, Gf[4 (q cells + c) ..] for q = 0..3 holding GradPlane entries 4q..4q+3 (the last
// group (mu_eff, k_eff, theta, 0)).  Every 16-byte access of a wavefront is contiguous.
constexpr int G32_GROUPS = 4;

struct Ptrs {
  double* p[WMX_MAX_BLOCKS];
};

struct BlockDev {
  int n[3];
  int N[3];
  int ng;
  int case_id;
  long long cells;          // N0 N1 N2
  const double* vol;
  const double* xc;
  const double* S[3];
  const double* xf[3];
  double* P;                // NP planes (FP64 form)
  double* G;                // NGR planes (FP64 form)
  float* Pf;                // NP planes (FP32-bulk form; same workspace region as P)
  float* Gf;                // NGR planes (FP32-bulk form)
  float* Sf[3];             // FP32-bulk form: float copies of S[d] (gradient kernel)
  float* volf;              // FP32-bulk form: float copy of vol
  double* Sb;               // CFL metric planes (NSB_PLANES: the cell means of the face vectors
                            // (d, x)), or nullptr: taken from S
  int sb_zero;              // bit i: plane i of Sb is +0.0 in every interior cell (an
                            // axis-aligned direction), so the dt takes 0.0 without loading it
  double* wall[6];          // WALL_PLANES planes of (n_t1, n_t2) per WallModel face
  double* sponge_mean;      // (5, n0, n1)
  int* sponge_row;          // (n0, n1): 1 if some cell of the row has a sponge weight > 0
  double* shell;            // (5, shell cells) saved ghost values
  long long shell_cells;
  int zfuse;                // bit s: the k ghosts of side s are the sweep's last phase, which
                            // the primitive kernels of a step evaluate themselves
  // FP32-bulk gauge of the block's case: p' = p - p_ref, T' = T - T_gauge,
  // e' = e_int - e_ref with e_ref = R T_gauge / (gamma - 1)
  double p_ref, T_gauge, e_ref;
  wmx_face_desc faces[6];
};

// Learned wall model handed to the wall kernels; w == nullptr selects the Newton model.
struct WallNetDev {
  const unsigned short* w;  // BF16 W2 [64][64] | W3 [2][64]
  const float* f;           // FP32 W1 [64][3] | b1 [64] | b2 [64] | b3 [2]
  float xc[3], xs[3];       // input normalisation x' = (x - xc) / xs
  float yc[2], ys[2];       // output de-normalisation log y = y' ys + yc
  double lo[3], hi[3];      // trained range of (log Re_h, h/Delta, T_h/T_w)
  int n_polish;             // Newton steps from the network's u+
  double tol;               // fallback when |F(u+)| / u+ exceeds it
  unsigned int* nfb;        // optional fallback counter
};

// host-evaluated constants of the FP32-bulk record store (rec32_store)
struct RecC {
  double iR, igm1;          // 1 / R, 1 / (gamma - 1)
  float iTref, muc, S;      // 1 / T_ref, mu_ref (T_ref + S_suth), S_suth
};

struct NumDev {
  GasC gas;
  RecC rc;
  WallModelC wm;
  int scheme, central, order, limiter;
  double th0, th1;
  int sgs;
  double sgs_c;
  double cfl;
  int n_sponge;
  double sponge[4][5];
  int prec;                 // WMX_PREC_FP64 / WMX_PREC_FP32_BULK (gauge: BlockDev)
  WallNetDev net;
};

constexpr int FLUX2_TILE_J = 15;   // cell rows of a flux-kernel tile (flux.hip)
// i cells per thread of the gradient kernel (rhs.hip): 4 in FP32, 1 in FP64
constexpr int grad_run(size_t rec_bytes) { return rec_bytes == 4 ? 4 : 1; }

struct GhostOp {
  int blk;       // destination block
  int axis, side;
  int count;     // ng * N_t1 * N_t2 ghost entries (cells or nodes)
};

struct WallOp {
  int blk, axis, side, count;
};

struct InflowDev {
  int n_tab, n_modes;
  int coop;                 // set by the launcher: phases evaluated cooperatively in LDS
  const double* table;
  const double* modes;
  double U_inf, L, rho_inf, p_inf;
};

// host-side launchers (one per kernel family; defined beside the kernels)
hipError_t launch_fill_nodes(const BlockDev* d_blks, const GhostOp* d_ops, int nops, int maxcount,
                             Ptrs X, hipStream_t s);
hipError_t launch_fill_cells(const BlockDev* d_blks, const GhostOp* d_ops, int nops, int maxcount,
                             Ptrs q, GasC gas, hipStream_t s);
hipError_t launch_inflow(const BlockDev* d_blks, int blk, int N1, int N2, int ng, InflowDev in,
                         Ptrs q, GasC gas, double time, const double* d_dt, double coef,
                         hipStream_t s);
hipError_t launch_to_f32(const double* a, float* b, long long n, hipStream_t s);
constexpr int NSB_PLANES = 9;
hipError_t launch_cell_sb(const BlockDev& hb, unsigned* d_nz, hipStream_t s);
hipError_t launch_geometry(const double* X, const int* N, double* vol, double* xc, double* const* S,
                           double* const* xf, hipStream_t s);
// d_part != nullptr: the launch also writes the CFL partial of each workgroup (step fusion)
// sponge row means taken by head workgroups of a primitive launch (pre == nullptr: none):
// pre[b] = first head workgroup of block b (32 rows each), nwg = their total
// FP32-bulk plans with fused gradients also give the Newton wall faces to head workgroups
// of the primitive launch, after the sponge ones: wall_wgx workgroups per wall op.
struct SpongeHead {
  const int* pre = nullptr;
  int nwg = 0;
  const WallOp* wall = nullptr;
  int wall_wgx = 0, nwall = 0;
  int head = 0;   // set by the launcher
};
// zfuse: the k ghosts flagged in BlockDev::zfuse are evaluated from the sweep (their fill
// launch is skipped) instead of read from q
hipError_t launch_prim(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                       const NumDev& nd, double* d_part, int zfuse, SpongeHead sh, hipStream_t s);
// per-case dt mins taken inside the gradient launch (out == nullptr: none)
struct GradDt {
  int ncases = 0;
  const int* list = nullptr;       // per-case workgroup list (nullptr: identity)
  const int* off = nullptr;        // per-case offsets into the list
  const double* part = nullptr;    // CFL partials
  double* out = nullptr;           // dt per case
};
// FP32-bulk plans: d_wall != nullptr also evaluates those Newton wall faces in the launch,
// gdt.out != nullptr the per-case dt mins
hipError_t launch_grad(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, const NumDev& nd,
                       const WallOp* d_wall, int nwall_ops, int wall_max, Ptrs q, const GradDt& gdt,
                       hipStream_t s);
hipError_t launch_wall(const BlockDev* d_blks, const WallOp* d_ops, int nops, int maxcount, Ptrs q,
                       const NumDev& nd, hipStream_t s);
hipError_t launch_flux(int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles, int ci_len,
                       const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt,
                       double dt, hipStream_t s);
// the fused FP32-bulk step kernel (flux3.h): gradient records evaluated in the march
hipError_t launch_flux3(int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles, int ci_len,
                        const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt, double dt,
                        hipStream_t s);
hipError_t launch_shell_save(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             hipStream_t s);
hipError_t launch_shell_finish(int mode, const BlockDev* d_blks, const int* d_pre, int nblk, int nwg,
                               Ptrs q0, Ptrs out, const double* d_dt, double dt, hipStream_t s);
hipError_t launch_dt_partial(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             const NumDev& nd, double* d_part, hipStream_t s);
constexpr int INT_CELLS = 4;      // interior cells per thread of the dt / mass-flux partials
constexpr int NSTAT = 12;          // plane statistics: rho u v w uu vv ww uv uw vw T p
constexpr int PLANE_CHUNKS = 16;   // workgroups per j plane in the plane statistics
hipError_t launch_plane_stats(const BlockDev* d_blks, const BlockDev& hb, int blk, Ptrs q,
                              const GasC& gas, double* d_part, double* d_acc, hipStream_t s);
hipError_t launch_mf_partial(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                             const int* d_case_dir, double* d_part, hipStream_t s);
hipError_t launch_case_final(int mode, int ncases, const int* d_case_wg, const int* d_case_off,
                             const double* d_part, const double* d_target, double* d_out,
                             hipStream_t s);
hipError_t launch_mf_apply(const BlockDev* d_blks, const int* d_pre, int nblk, int nwg, Ptrs q,
                           const int* d_case_dir, const double* d_delta, hipStream_t s);
hipError_t launch_wm_eval(const NumDev& nd, int mode, long long n, const double* a, const double* b,
                          const double* c, const double* d, double T_wall, double* o0, double* o1,
                          hipStream_t s);
// learned wall model (mlp.hip)
hipError_t launch_wm_dataset(const NumDev& nd, long long n, unsigned long long seed, float* X,
                             float* Y, double* raw, hipStream_t s);
hipError_t launch_mlp_train(float* P, const float* X, const float* Y, long long off, int batch,
                            float* gpart, float* lpart, float* gout, float* m, float* v, int step,
                            float lr, int update, hipStream_t s);
hipError_t launch_mlp_pack(const float* P, unsigned short* W, float* F, hipStream_t s);
hipError_t launch_mlp_forward32(const float* P, const float* X, long long n, float* Y, hipStream_t s);
hipError_t launch_wm_net_eval(const NumDev& nd, int mode, long long n, const double* h,
                              const double* hdel, const double* U, const double* Th,
                              const double* ph, double T_wall, double* tau, double* qw,
                              double* uplus, hipStream_t s);
hipError_t launch_mfma_probe(const float* A, const float* B, float* D, hipStream_t s);
// Scheme configurations compiled as flux-kernel instances of their own (the face flux then
// carries no runtime dispatch); SC_ANY reads everything from NumDev.
enum { SC_ANY = 0, SC_HYB_KEP4_ALBADA = 1, SC_KEP2 = 2, SC_KEP4 = 3 };

__host__ __device__ inline int scheme_config(const NumDev& nd) {
  if (nd.scheme == WMX_SCHEME_HYBRID && nd.central == WMX_SCHEME_KEP && nd.order != 2 && nd.limiter == 0)
    return SC_HYB_KEP4_ALBADA;
  if (nd.scheme == WMX_SCHEME_KEP) return nd.order == 2 ? SC_KEP2 : SC_KEP4;
  return SC_ANY;
}

// One RK stage of an FP32-bulk plan as a single cooperative (persistent) launch
// (stage.hip): the sweep's phases, the primitive records (+ CFL partials at stage 1), the
// gradients with their head work (per-case dt min, Newton wall faces) and the flux tiles,
// separated by grid-wide barriers.  What does not change between calls lives in device
// memory (StageStatic, uploaded by the plan); the state pointers and the stage's own values
// come as kernel arguments (StageCall), and the launch's workgroup 0 copies the pointer
// tables to StageStatic::pbuf before the first barrier, so that the bodies index them in
// global memory.
struct StageStatic {
  const BlockDev* blks;
  NumDev nd;
  const GhostOp* ops;          // sweep phases: ops[fill_off[ph] .. fill_off[ph + 1])
  const int* fill_off;
  const int* fill_max;         // cells of the phase's largest op
  const int* pre_all;          // primitive workgroups per block (prefix)
  int nblk, nprim;
  const int* pre_box;          // gradient workgroups per block (prefix)
  int ngrad;
  const WallOp* wall_ops;      // Newton wall faces: in the gradient head (grad_wall_wgx
  int nwall_ops;               // workgroups per op) or, fused plans, in the primitive head
  int grad_wall_wgx, prim_wall_wgx, prim_head;
  const int* dt_list;          // per-case workgroup lists of the CFL partials
  const int* dt_off;
  const int4* tiles;
  int ntiles, ci_len;
  Ptrs* pbuf;                  // [3]: qin, q0, out of the running launch
};
struct StageCall {
  Ptrs qin, q0, out;
  int nfill, zfuse, head, ndt;
  double* part;                // CFL partials (stage 1) or nullptr
  double* dt_out;              // per-case dt (stage 1) or nullptr
  const double* d_dt;
  double dt_host;
};
// fused: the fused kernel form (gradient records in the flux march, wall faces in the
// primitive head).  grid: workgroups of the persistent grid (0: the co-resident maximum, a
// multiple of 8, returned in *grid_used)
hipError_t launch_stage(int mode, int sc, bool fused, const StageStatic* d_st, const StageCall& c,
                        int* grid_used, hipStream_t s);

hipError_t launch_scheme_flux(const NumDev& nd, const int* n, int d, const double* w, const double* S,
                              const double* theta, double* F, hipStream_t s);
hipError_t launch_sgs_eval(const NumDev& nd, long long n, const double* g, const double* delta,
                           const double* rho, double* mut, hipStream_t s);
hipError_t launch_wall_flux(const NumDev& nd, const BlockDev& B, int a, int side, const double* w,
                            const double* T, double* F, double* tau, double* qw, hipStream_t s);
int mlp_nparam();
int mlp_nbf16();
int mlp_nf32();
hipError_t launch_sponge_rows(const BlockDev* d_blks, int blk, long long rows, const NumDev& nd,
                              hipStream_t s);
hipError_t launch_sponge_mean(const BlockDev* d_blks, int blk, long long rows, Ptrs q,
                              hipStream_t s);


// Global-memory views of device pointers.  Loads and stores through them compile to
// global_load / global_store instead of flat ones; a flat access may alias LDS, so the
// compiler would otherwise complete each one before the next LDS access (a full memory
// round trip per element in the LDS-staging loops).
template <class T>
using gconst = const __attribute__((address_space(1))) T*;
template <class T>
using gmut = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gconst<T> gmem(const T* p) { return (gconst<T>)p; }
template <class T>
__device__ __forceinline__ gmut<T> gmem_w(T* p) { return (gmut<T>)p; }

// Load / store at element `i` of an array whose base pointer is wave-uniform, with the
// byte offset formed in 32 bits: the access compiles to the scalar-base global form
// (64-bit base in SGPRs, one 32-bit offset VGPR).  The caller guarantees i * sizeof(T)
// < 2^32 (plans reject blocks whose planes would exceed that).
template <class T>
__device__ __forceinline__ T ldu(const T* base, unsigned i) {
  return *(gconst<T>)((const __attribute__((address_space(1))) char*)base + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stu(T* base, unsigned i, T v) {
  *(gmut<T>)((__attribute__((address_space(1))) char*)base + i * (unsigned)sizeof(T)) = v;
}

// Workgroup w of a launch runs on XCD w % 8.  xcd_span maps it to a logical index so
// that each XCD takes one contiguous span of the launch (sizes differ by at most one):
// workgroups that share halo data then share an L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_span(int w, int n) {
  const int q = n >> 3, r = n & 7, x = w & 7;
  return x * q + min(x, r) + (w >> 3);
}

// 16- and 8-byte record accesses (global address space; rec 16- / 8-byte aligned: the
// workspace is 256-byte aligned and the record strides are multiples of 4 floats)
typedef float wmx_f4 __attribute__((ext_vector_type(4)));
typedef float wmx_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 ld4(const float* rec) {
  const wmx_f4 v = *(const __attribute__((address_space(1))) wmx_f4*)__builtin_assume_aligned(rec, 16);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ld2(const float* rec) {
  const wmx_f2 v = *(const __attribute__((address_space(1))) wmx_f2*)__builtin_assume_aligned(rec, 8);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st4(float* rec, float4 v) {
  wmx_f4 w = {v.x, v.y, v.z, v.w};
  *(__attribute__((address_space(1))) wmx_f4*)__builtin_assume_aligned(rec, 16) = w;
}

// FP32-bulk primitive record of a cell from its conserved state q, with the block gauge
// (p' = p - p_ref, T' = T - T_gauge, e' = e_int - e_ref): cons_to_prim (reference
// gas.py:28-33) in FP64 with one reciprocal of rho, and Sutherland's mu (gas.py:24-26)
// from the float temperature (the record holds it as a float), stored as the two 16-byte
// groups (u, v, w, T') and (rho, p', mu, e').  The arithmetic is not contracted (callers are
// built with contraction off or, where fast, honour the pragma), so every caller stores the
// same bits.
__device__ __forceinline__ void rec32_store(float* Pb, size_t cells, size_t c, const NumDev& nd, const double* q,
                                            double p_ref, double T_gauge, double e_ref) {
#pragma clang fp contract(off)
  const RecC& rc = nd.rc;
  const double r = q[0];
  const double ir = 1.0 / r;
  const double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
  const double k = (u * u + v * v) + w * w;
  const double p = (nd.gas.gamma - 1.0) * (q[4] - (0.5 * r) * k);
  const double pr = p * ir;                      // p / rho = R T = (gamma - 1) e_int
  const double T = pr * rc.iR;
  const float Tf = (float)T;
  const float t = Tf * rc.iTref;
  const float mu = (rc.muc * (t * __builtin_amdgcn_sqrtf(t))) * __builtin_amdgcn_rcpf(Tf + rc.S);
  st4(Pb + 4 * c, make_float4((float)u, (float)v, (float)w, (float)(T - T_gauge)));
  st4(Pb + 4 * (cells + c), make_float4((float)r, (float)(p - p_ref), mu, (float)(pr * rc.igm1 - e_ref)));
}

// SmoothRampSponge weight at a cell centre x (smooth_ramp.py:212-221); pr = (x_start_l,
// x_end_l, strength, U_inf, L)
__host__ __device__ inline double sponge_sigma(const double* pr, double xcv) {
  const double x_l = xcv / pr[4];
  const double eta = fmin(fmax((x_l - pr[0]) / (pr[1] - pr[0]), 0.0), 1.0);
  const double ramp = (eta * eta * eta) * (10.0 + eta * (-15.0 + 6.0 * eta));
  return ((pr[2] * pr[3]) / pr[4]) * ramp;
}

// tangential axes of a face normal to `axis`, increasing order
WMX_HD void tangents(int axis, int* t1, int* t2) {
  *t1 = axis == 0 ? 1 : 0;
  *t2 = axis == 2 ? 1 : 2;
}

WMX_HD long long cidx(const int* N, int I, int J, int K) {
  return ((long long)I * N[1] + J) * N[2] + K;
}

// face-vector array S[d] has N_d + 1 entries along d
WMX_HD long long fidx(const int* N, int d, int I, int J, int K) {
  int M1 = N[1] + (d == 1), M2 = N[2] + (d == 2);
  return ((long long)I * M1 + J) * M2 + K;
}
WMX_HD long long fcount(const int* N, int d) {
  return (long long)(N[0] + (d == 0)) * (N[1] + (d == 1)) * (N[2] + (d == 2));
}

// ghost shell cell s -> (I, J, K): i-slabs (full j,k), j-slabs (interior i), k-slabs
// (interior i, j); k fastest within each.  32-bit arithmetic: a plan's blocks hold fewer
// than 2^29 cells (each FP64 plane stays below 4 GiB), and the shell is part of the block.
WMX_HD void shell_cell(unsigned s, const int* N, const int* n, int ng, int* I, int* J, int* K) {
  const unsigned N1 = (unsigned)N[1], N2 = (unsigned)N[2], ung = (unsigned)ng;
  const unsigned a = 2u * ung * N1 * N2;
  if (s < a) {
    const unsigned l = s / (N1 * N2), r = s - l * (N1 * N2);
    const unsigned j = r / N2;
    *I = l < ung ? (int)l : (int)(n[0] + l);
    *J = (int)j;
    *K = (int)(r - j * N2);
    return;
  }
  s -= a;
  const unsigned per_j = (unsigned)n[0] * N2;
  if (s < 2u * ung * per_j) {
    const unsigned l = s / per_j, r = s - l * per_j;
    const unsigned i = r / N2;
    *I = ng + (int)i;
    *K = (int)(r - i * N2);
    *J = l < ung ? (int)l : (int)(n[1] + l);
    return;
  }
  s -= 2u * ung * per_j;
  // k-slabs: interior (I, J) row, then its low side's ng layers, then its high side's.  In
  // memory a row's high ghosts and the next row's low ghosts are adjacent, so a wavefront
  // touches runs of 2 ng cells (one cell per row per lane, or one side per row, would touch
  // one or ng).
  const unsigned row = s / (2u * ung);
  const unsigned t = s - row * (2u * ung);
  const int side = t < ung ? 0 : 1;
  const int m = (int)(t - (side ? ung : 0u));
  const unsigned i = row / (unsigned)n[1];
  *I = ng + (int)i;
  *J = ng + (int)(row - i * (unsigned)n[1]);
  *K = side == 0 ? m : ng + n[2] + m;
}

// k ghost of shell cell (I, J, K) whose side the step's primitive launches evaluate (the
// sweep's last phase, BlockDev::zfuse): a step's first stage leaves it unwritten in q0
WMX_HD bool shell_kept(const int* n, int ng, int zfuse, int K) {
  if (!zfuse) return false;
  if (K < ng) return zfuse & 1;
  if (K >= ng + n[2]) return (zfuse >> 1) & 1;
  return false;
}

WMX_HD long long shell_count(const int* N, const int* n, int ng) {
  return 2LL * ng * ((long long)N[1] * N[2] + (long long)n[0] * N[2] + (long long)n[0] * n[1]);
}

// numpy's pairwise summation (loops_utils pairwise_sum) of elem(0..n-1), written
// iteratively: runs of <= 128 are summed with 8 interleaved accumulators and longer
// runs are split at n/2 rounded down to a multiple of 8.
template <class F>
__host__ __device__ double np_pairwise_leaf(const F& elem, int off, int n) {
  if (n < 8) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r = r + elem(off + i);
    return r;
  }
  double r[8];
  for (int u = 0; u < 8; ++u) r[u] = elem(off + u);
  int i = 8;
  for (; i < n - (n % 8); i += 8)
    for (int u = 0; u < 8; ++u) r[u] = r[u] + elem(off + i + u);
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + elem(off + i);
  return res;
}

template <class F>
__host__ __device__ double np_pairwise(const F& elem, int n) {
  int w_off[64], w_n[64], w_t[64];
  double vals[40];
  int nw = 0, nv = 0;
  w_off[0] = 0; w_n[0] = n; w_t[0] = 0; nw = 1;
  while (nw > 0) {
    --nw;
    const int off = w_off[nw], len = w_n[nw], t = w_t[nw];
    if (t == 1) {
      const double b = vals[--nv];
      const double a = vals[--nv];
      vals[nv++] = a + b;
    } else if (len <= 128) {
      vals[nv++] = np_pairwise_leaf(elem, off, len);
    } else {
      int h = len / 2;
      h -= h % 8;
      w_t[nw] = 1; w_off[nw] = 0; w_n[nw] = 0; ++nw;
      w_t[nw] = 0; w_off[nw] = off + h; w_n[nw] = len - h; ++nw;
      w_t[nw] = 0; w_off[nw] = off; w_n[nw] = h; ++nw;
    }
  }
  return vals[0];
}

// numpy add.reduce of a run: the first element, then the pairwise sum of the rest
template <class F>
__host__ __device__ double np_sum(const F& elem, int n) {
  if (n <= 0) return 0.0;
  struct Shift {
    const F& e;
    __host__ __device__ double operator()(int i) const { return e(i + 1); }
  } rest{elem};
  return elem(0) + np_pairwise(rest, n - 1);
}

}  // namespace wmx

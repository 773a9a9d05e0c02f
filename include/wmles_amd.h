/*
 * wmles_amd.h -- C ABI of libwmles_amd.so, the MI355X (gfx950) explicit-step hot path
 * of a compressible wall-modelled LES finite-volume solver.
 *
 * The reference exposes this path as Python call signatures only (no FFI); each
 * entry point below replaces one of them, cited as reference file:line.  Python
 * binds these symbols with ctypes (wmles_jax_amd/_lib.py).
 *
 * Conventions
 *  - Every pointer argument named q / r / X / vol / xc / S / xf / d_* is DEVICE
 *    memory allocated and owned by the caller.  The library borrows it for the call
 *    (geometry pointers: for the plan's lifetime).  The plan owns only its
 *    metadata (ghost-fill op lists, tile tables), freed by wmx_plan_destroy.
 *  - State arrays use the reference layout: one float64 array per block of shape
 *    (5, Ni, Nj, Nk) C-order, Ni = ni + 2 ng, q = [rho, rho u, rho v, rho w, rho E]
 *    (reference wmles/gas.py:3-4, wmles/mesh.py:3-5).
 *  - Return value: 0 on success, non-zero on error; wmx_last_error() returns the
 *    message (thread-local).  No C++ exception crosses the ABI.
 *  - `stream` is a hipStream_t passed as void* (0 = default stream).  All work is
 *    enqueued on it; nothing synchronises the host unless stated.
 *  - A plan is not re-entrant: calls on one plan must not overlap.
 */
#ifndef WMLES_AMD_H
#define WMLES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WMX_MAX_BLOCKS 64

/* face kinds: reference wmles/mesh.py:77-85 (Interface) and wmles/bc.py:38-166 */
enum {
  WMX_FACE_INTERFACE = 0,
  WMX_BC_SLIP = 1,           /* SlipWall                  bc.py:38-47   */
  WMX_BC_NOSLIP = 2,         /* NoSlipWall                bc.py:50-64   */
  WMX_BC_DIRICHLET = 3,      /* Dirichlet                 bc.py:67-77   */
  WMX_BC_FARFIELD = 4,       /* CharacteristicFarfield    bc.py:80-115  */
  WMX_BC_EXTRAPOLATE = 5,    /* Extrapolate               bc.py:118-125 */
  WMX_BC_WALLMODEL = 6,      /* WallModel (slip fill + flux override) bc.py:128-166 */
  WMX_BC_INFLOW = 7,         /* SmoothRampInflow          smooth_ramp.py:116-197 */
  WMX_BC_EXTERNAL = 8        /* a BC the caller fills itself (a user fill(q, blk, geom, gas,
                                axis, side) without a HIP descriptor, bc.py:1-8): the sweep
                                gives it a phase of its own and leaves it to the caller */
};

/* inviscid schemes: reference wmles/schemes.py */
enum { WMX_SCHEME_KEP = 0, WMX_SCHEME_KEEP = 1, WMX_SCHEME_MUSCLHLLC = 2, WMX_SCHEME_HYBRID = 3 };
enum { WMX_LIMITER_ALBADA = 0, WMX_LIMITER_MINMOD = 1 };
/* SGS models: reference wmles/sgs.py:45-50 */
enum { WMX_SGS_NONE = 0, WMX_SGS_SMAGORINSKY = 1, WMX_SGS_WALE = 2, WMX_SGS_VREMAN = 3 };
/* arithmetic forms */
enum { WMX_PREC_FP64 = 0, WMX_PREC_FP32_BULK = 1 };

typedef struct wmx_face_desc {
  int32_t kind;
  int32_t nb;              /* interface: source block index                    */
  int32_t src_axis;        /* interface: (axis, side) of the face on the source */
  int32_t src_side;
  int32_t mapping[3];      /* interface: destination axis t -> source axis     */
  int32_t flip[3];         /* interface: reverse destination tangential axis t */
  double shift[3];         /* interface: x_dst = x_src + shift (node exchange)  */
  double prim[5];          /* Dirichlet / far field: [rho, u, v, w, p]          */
  double T_wall;           /* NoSlip / WallModel wall temperature; < 0: adiabatic */
  int32_t n_match;         /* WallModel: matching cell index off the wall (>= 1) */
  int32_t inflow_id;       /* Inflow: index into the inflow descriptor list     */
} wmx_face_desc;

typedef struct wmx_block_desc {
  int32_t n[3];            /* interior cells ni, nj, nk                          */
  int32_t case_id;         /* ensemble case of this block (0 for a single case)  */
  wmx_face_desc faces[6];  /* imin, imax, jmin, jmax, kmin, kmax                 */
  /* geometry (device, reference layout wmles/mesh.py:38-43), borrowed for the
     plan's lifetime:  vol (Ni,Nj,Nk); xc (3,Ni,Nj,Nk); S[d], xf[d] (3, ..N_d+1..) */
  const double* vol;
  const double* xc;
  const double* S[3];
  const double* xf[3];
} wmx_block_desc;

typedef struct wmx_numerics_desc {
  /* gas: reference wmles/gas.py:10-18 */
  double gamma, R, Pr, Pr_t, mu_ref, T_ref, S_suth;
  /* scheme: reference wmles/schemes.py:57-166 */
  int32_t scheme;          /* WMX_SCHEME_*                                      */
  int32_t central;         /* Hybrid central part: WMX_SCHEME_KEP or _KEEP      */
  int32_t order;           /* KEP/KEEP (or Hybrid central) order: 2 or 4        */
  int32_t limiter;         /* MUSCL limiter: WMX_LIMITER_*                      */
  double th0, th1;         /* Hybrid ramp                                       */
  /* SGS: reference wmles/sgs.py */
  int32_t sgs;             /* WMX_SGS_*                                         */
  double sgs_c;            /* Cs / Cw / Cs(Vreman)                              */
  /* wall model: reference wmles/wall_model.py:15-17; the four constants are the
     host-evaluated exp(-kappa B), Pr^(1/3), Pr^3 and Kader's beta (wall_model.py:21,
     35, 45, 48) so that they round exactly as the reference's host arithmetic. */
  double kappa, B;
  int32_t n_iter;
  double wm_ekb, wm_pr13, wm_pr3, wm_beta;
  int32_t bc_passes;       /* reference Solver(bc_passes=2)                     */
  double cfl;              /* reference Solver(cfl=0.8)                         */
  int32_t precision;       /* WMX_PREC_*                                        */
  /* sponge sources (reference smooth_ramp.py:200-221), applied in list order */
  int32_t n_sponge;
  double sponge[4][5];     /* x_start_l, x_end_l, strength, U_inf, L            */
  /* ensemble: per-case constant-mass-flux constraint (reference solver.py:16-36) */
  int32_t n_cases;
  int32_t mass_flux_dir[16];   /* -1: no constraint for that case               */
  double mass_flux_target[16];
} wmx_numerics_desc;

typedef struct wmx_inflow_desc {
  /* reference smooth_ramp.py:116-197 (SmoothRampInflow) */
  int32_t n_tab;           /* rows of the profile table (>= 1)                  */
  const double* table;     /* device (n_tab, 7): y/L, U, V, uu, vv, ww, uv      */
  int32_t n_modes;         /* >= 1; wmx_plan_create rejects 0 or a NULL array   */
  const double* modes;     /* device (3 components, 4, n_modes): kx, ky, kz, phase */
  double U_inf, L, rho_inf, p_inf;
} wmx_inflow_desc;

/* Learned wall model (no reference code: it plugs in at the wall-model callable
   boundary model(gas, h, U, T_h, p_h, T_wall), reference wmles/wall_model.py:3-9,
   called at wmles/bc.py:161).  MLP 3 -> 64 -> 64 -> 2, hidden activation a / sqrt(1 + a^2);
   inputs log Re_h, h/Delta, T_h/T_w; outputs log u+ and log T+ (normalised).
   FP32 master parameters (wmx_mlp_sizes: nparam) are ordered
   W1[64][3] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]; wmx_mlp_pack derives the
   inference copies w_bf16 = W2 | W3 (BF16 bits) and w_fp32 = W1 | b1 | b2 | b3. */
typedef struct wmx_wall_net_desc {
  const uint16_t* w_bf16;   /* device, nbf16 entries                                */
  const float* w_fp32;      /* device, nfp32 entries                                */
  float x_center[3], x_scale[3];   /* x' = (x - center) / scale                    */
  float y_center[2], y_scale[2];   /* log y = y' scale + center                    */
  double lo[3], hi[3];      /* trained range of (log Re_h, h/Delta, T_h/T_w); a point
                               outside the log Re_h or T_h/T_w range falls back to
                               the full Newton solve; h/Delta is clamped (the
                               equilibrium law does not depend on it)             */
  int32_t n_polish;         /* Newton steps from the network's u+                   */
  double tol;               /* fall back when |F(u+)| / u+ > tol after the polish   */
  uint32_t* fallback_count; /* optional device counter, incremented per fallback    */
} wmx_wall_net_desc;

typedef struct wmx_plan wmx_plan;

/* Last error message of the calling thread ("" if none). */
const char* wmx_last_error(void);
/* Library build description (arch, flags). */
const char* wmx_build_info(void);

/* Build a plan: topology, BC parameters, numerics, tile tables and ghost-fill
   phases.  Replaces the state the reference keeps in Mesh/Solver
   (wmles/mesh.py:196-242, wmles/solver.py:40-51). */
int wmx_plan_create(int32_t nblocks, const wmx_block_desc* blocks,
                    const wmx_numerics_desc* numerics, int32_t ninflow,
                    const wmx_inflow_desc* inflows, wmx_plan** out);
int wmx_plan_destroy(wmx_plan* plan);
/* Device workspace the plan needs (bytes, 256-aligned); bind it once. */
int wmx_plan_workspace_bytes(const wmx_plan* plan, size_t* bytes);
int wmx_plan_bind_workspace(wmx_plan* plan, void* d_workspace, size_t bytes);
/* Number of WallModel faces and their (block, axis, side, n_t1, n_t2) records. */
int wmx_plan_wall_faces(const wmx_plan* plan, int32_t* count, int32_t* records /* 5*count or NULL */);
/* Ghost-fill phases after dependency grouping (diagnostic). */
int wmx_plan_fill_phases(const wmx_plan* plan, int32_t* nphases);
/* Sweep passes executed: bc_passes minus the trailing passes shown (by a symbolic run
   of the sweep at plan creation) to reproduce the previous pass bit for bit. */
int wmx_plan_fill_passes(const wmx_plan* plan, int32_t* npasses);

/* FP32-bulk gauge (no reference counterpart: precision form of this build).  The
   FP32 records hold p - p_ref, T - T_ref and e_int - R T_ref / (gamma - 1); the
   fluxes add the shift's contribution back in FP64.  Values near the flow's mean
   pressure and temperature keep the low-Mach cancellation in FP64.  Ignored by the
   FP64 plain form.  wmx_plan_set_gauge sets every case of the plan;
   wmx_plan_set_case_gauge the blocks of one ensemble case (case_id -1: all), so
   cases with different mean states each keep their own FP32 range.  Both synchronise
   the device when a workspace is bound (the values live in the device block
   descriptors, which captured step graphs read at replay). */
int wmx_plan_set_gauge(wmx_plan* plan, double p_ref, double T_ref);
int wmx_plan_set_case_gauge(wmx_plan* plan, int32_t case_id, double p_ref, double T_ref);

/* Kernel-class event timing (measurement aid, no reference counterpart).  Classes:
   0 flux+RK, 1 gradient/SGS, 2 primitives, 3 ghost fill, 4 wall model, 5 CFL dt,
   6 mass flux, 7 ghost shell, 8 inflow fill, 9 sponge mean.  wmx_plan_profile(plan, mask)
   brackets every later
   launch of a class whose bit is set in `mask` with hipEvents on the launch stream
   (mask 0 stops and clears); wmx_plan_profile_read synchronises the recorded events
   and returns per-class summed milliseconds and launch counts (10 entries each). */
int wmx_plan_profile(wmx_plan* plan, int32_t mask);
int wmx_plan_profile_read(wmx_plan* plan, double* ms, int32_t* counts);

/* Setup: ghost nodes by node exchange / point reflection, `node_passes` sweeps.
   X[b]: device (3, Ni+1, Nj+1, Nk+1) with interior nodes set.
   Replaces wmles/mesh.py:213-232 (Mesh.finalize ghost-node loop, _reflect_nodes). */
int wmx_fill_nodes(int32_t nblocks, const wmx_block_desc* blocks, double* const* X,
                   int32_t ng, int32_t node_passes, void* stream);
/* Setup: face vectors, face centres, volumes and centroids of one block from its
   ghost-extended nodes.  Replaces wmles/mesh.py:53-73 (compute_geometry). */
int wmx_compute_geometry(const double* X, const int32_t n_ghosted[3], double* vol, double* xc,
                         double* S0, double* S1, double* S2,
                         double* xf0, double* xf1, double* xf2, void* stream);

/* Ghost sweep in place (interior untouched).  Replaces Solver.apply_bcs
   (wmles/solver.py:59-73).  Plans with WMX_BC_EXTERNAL faces refuse it (and wmx_rhs,
   wmx_advance, wmx_step): their caller runs the sweep phase by phase. */
int wmx_apply_bcs(wmx_plan* plan, double* const* q, double time, void* stream);
/* The sweep's phases in execution order: records[4 * i] = 1 when phase i is the fill of
   a WMX_BC_EXTERNAL face (then records[4 i + 1..3] = block, axis, side), else 0 (and -1).
   records may be NULL to get the count.  With the reference's sequential order
   (solver.py:59-73) a caller fills the external faces itself between
   wmx_apply_bcs_phases calls. */
int wmx_plan_sweep_phases(const wmx_plan* plan, int32_t* nphases, int32_t* records);
/* Runs phases [first, last) of the sweep in place; the range must hold no external
   phase. */
int wmx_apply_bcs_phases(wmx_plan* plan, double* const* q, int32_t first, int32_t last, double time,
                         void* stream);
/* R = RHS(q) with zero ghosts, for a q whose ghosts the caller has filled (no sweep; q is
   not modified).  The RHS part of Solver._rhs after apply_bcs (wmles/solver.py:97-106);
   the callers' own sources are added by the caller. */
int wmx_rhs_filled(wmx_plan* plan, double* const* q, double* const* r, void* stream);
/* R = RHS(q) with zero ghosts; q's ghost cells are overwritten by the sweep.
   Replaces Solver.rhs / _rhs (wmles/solver.py:97-106,159-160). */
int wmx_rhs(wmx_plan* plan, double* const* q, double* const* r, double time, void* stream);
/* One SSP-RK3 step + constraints: qout = step(q0).  The three stages run in place
   in qout (qout must not alias q0).  q0 is returned unchanged (its ghosts are saved
   to the workspace and restored); qout's ghost cells carry q0's ghost values
   through the stage combinations with R = 0 there, as in the reference.  dt is
   read from d_dt[case] (device) when d_dt is non-NULL, else the host value dt.
   Replaces Solver.advance / _advance_staged (wmles/solver.py:109-133,156-157). */
int wmx_advance(wmx_plan* plan, double* const* q0, double* const* qout, double dt,
                const double* d_dt, double time, void* stream);
/* The reference driver's step (examples/channel_wmles.py:124-125: max_dt then advance)
   with dt kept on the device: d_dt[case] <- CFL dt of q0, then qout = advance(q0).
   The dt partials come from the first stage's primitive launch (same values as
   wmx_max_dt, one pass over q0 fewer).
   use_graph != 0 (and a non-default stream, and no time-dependent inflow): the launch
   sequence is captured once per (q0, qout, d_dt, time) into a hipGraph (at most 4
   kept per plan) and replayed on later calls with the same pointers. */
int wmx_step(wmx_plan* plan, double* const* q0, double* const* qout, double* d_dt, double time,
             int32_t use_graph, void* stream);
/* Flux-kernel tiling (autotuner knob; no counterpart in the reference): ci_len
   i-planes per marching tile (0: chosen from a makespan model of the device's CU
   count, the default); order 0 (block/i/j/k) or 1 (contiguous spans interleaved
   across the 8 XCDs, the default). */
int wmx_plan_set_tiling(wmx_plan* plan, int32_t ci_len, int32_t order);
/* FP32-bulk kernel form (no reference counterpart): on != 0 (the default) the flux kernel
   evaluates the Green-Gauss gradients, SGS viscosity and Ducros sensor of each plane in
   LDS as it marches, so no gradient records go through HBM and the separate gradient launch
   keeps only its side work (per-case dt min, Newton wall faces); on == 0 runs the separate
   gradient kernel and the flux kernel that reads its records.  Persistent stages follow
   the same choice; FP64 plans ignore it. */
int wmx_plan_set_fused_gradients(wmx_plan* plan, int32_t on);
/* Persistent stages (the batching form's launch structure; no reference counterpart):
   on != 0 runs every RK stage of wmx_advance / wmx_step as ONE cooperative launch -- a
   grid of the kernel's co-resident maximum walks the stage's ghost-sweep phases, the
   primitive records, the gradients (with the per-case dt min and the Newton wall faces)
   and the flux tiles, with grid-wide barriers between them -- instead of one launch per
   kernel class; per step the ghost-shell save / restore and the mass-flux constraint stay
   separate launches.  Same work and same results (bit for bit) as the separate launches.
   FP32-bulk plans without time-dependent inflow, sponge sources, the learned wall model or
   caller-filled faces only (others are refused with an error); not captured into graphs
   (wmx_step's use_graph is ignored while it is on).  on == 0 restores the separate
   launches. */
int wmx_plan_set_persistent(wmx_plan* plan, int32_t on);
/* The automatic choice without a plan (host only): i-planes per tile and tile count
   for blocks of interior sizes dims[3 * nblocks] in the given form on cu_count CUs. */
int wmx_tile_model(int32_t nblocks, const int32_t* dims, int32_t precision, int32_t cu_count,
                   int32_t* ci_len, int32_t* ntiles);
/* The tiling in use: i-planes per tile, order, number of tiles. */
int wmx_plan_tiling(const wmx_plan* plan, int32_t* ci_len, int32_t* order, int32_t* ntiles);
/* CFL time step (per case when the plan is an ensemble) into d_dt[n_cases].
   Replaces Solver.max_dt / _dt (wmles/solver.py:135-154). */
int wmx_max_dt(wmx_plan* plan, double* const* q, double* d_dt, void* stream);
/* Allocation-free CFL time step (no counterpart in the reference: it is for callers that
   cannot allocate device memory -- code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  The same result as wmx_max_dt, bit for bit, with the
   per-workgroup CFL partials in a static device array of the library (one per GPU,
   WMX_NOALLOC_MAX_PARTIALS doubles = 512 KiB, held until the process ends) instead of
   the plan's workspace; the plan needs no workspace bound.  Makes no hipMalloc / hipFree.
   Plans whose interior cells need more partials than the array holds (more than
   WMX_NOALLOC_MAX_PARTIALS x 1024 interior cells, about 67 M) are rejected with an error.
   Every call on one GPU shares that GPU's array: calls that overlap on one GPU (other
   streams, other threads) race and must be serialised by the caller. */
#define WMX_NOALLOC_MAX_PARTIALS 65536
int wmx_max_dt_noalloc(wmx_plan* plan, double* const* q, double* d_dt, void* stream);
/* Constant-mass-flux constraint in place with the plan's per-case targets.
   Replaces ConstantMassFlux.__call__ (wmles/solver.py:22-36). */
int wmx_mass_flux(wmx_plan* plan, double* const* q, void* stream);
/* tau_w and q_w of every WallModel face, concatenated in wmx_plan_wall_faces
   order, each face (n_t1, n_t2).  Replaces Solver.wall_stats (wmles/solver.py:163-179). */
int wmx_wall_stats(wmx_plan* plan, double* const* q, double* d_tau, double* d_qw, void* stream);
/* Plane statistics of one block: adds, for every interior j plane, the (i, k) means of
   rho, u, v, w, uu, vv, ww, uv, uw, vw, T, p to d_acc[12][n_j] (FP64, device).
   Replaces the averaging of PlaneStats.accumulate (wmles/channel5200.py:139-157). */
int wmx_plane_stats(wmx_plan* plan, double* const* q, int32_t block, double* d_acc, void* stream);

/* Fill the ng ghost layers of one boundary (non-interface, non-inflow) face of one
   block in place; synchronises the stream.  The BC protocol's fill(q, blk, geom,
   gas, axis, side) (reference wmles/bc.py:31-35, _fill_mirror bc.py:14-25). */
int wmx_fill_face(const wmx_block_desc* block, int32_t axis, int32_t side,
                  const wmx_numerics_desc* numerics, double* q, double time, void* stream);
/* Equilibrium wall model on n points (reference wmles/wall_model.py:19-61).
   mode 0: out0 = u_tau(h=a, U=b, rho_w=c, mu_w=d)              (utau)
   mode 1: out0 = tau_w, out1 = q_w from (h=a, U=b, T_h=c, p_h=d), T_wall (< 0:
           adiabatic)                                           (__call__) */
int wmx_wall_model_eval(const wmx_numerics_desc* numerics, int32_t mode, int64_t n,
                        const double* a, const double* b, const double* c, const double* d,
                        double T_wall, double* out0, double* out1, void* stream);

/* ---- the reference's operator callables, evaluated in FP64 by the step's own device code ---- */
/* Inviscid flux incl. area of the numerics' scheme (KEP / KEEP of order 2 or 4,
   MUSCL-HLLC, Hybrid) through the n_d + 1 interior faces of direction d over the
   interior tangential cells: F (5, n0 + (d==0), n1 + (d==1), n2 + (d==2)) C order.
   w: (5, N0, N1, N2) ghosted primitives [rho, u, v, w, p] (ng = 3); S_d: the block's
   face vectors of direction d (3, .. N_d + 1 ..); theta: (N0, N1, N2) Ducros sensor
   (Hybrid only, else may be NULL).  Replaces scheme(blk, geom, w, gas, d, theta)
   (reference wmles/schemes.py:3-5, KEP 57-70, KEEP 72-98, MUSCLHLLC 140-152, Hybrid 155-166). */
int wmx_scheme_flux(const wmx_numerics_desc* numerics, const int32_t n[3], int32_t d, const double* w,
                    const double* S_d, const double* theta, double* F, void* stream);
/* Eddy viscosity mu_t = rho nu_t(g, delta) of the numerics' SGS model on n points:
   g (3, 3, n) with g[i][j] = du_i/dx_j, delta and rho (n).  Replaces the callable of
   make_sgs(name, **kw)(g, delta, rho) (reference wmles/sgs.py:6-50). */
int wmx_sgs_eval(const wmx_numerics_desc* numerics, int64_t n, const double* g, const double* delta,
                 const double* rho, double* mu_t, void* stream);
/* Wall-model flux of WallModel face (axis, side) of one block with the equilibrium model:
   F (5, n_t1, n_t2) incl. area and the sign of the face's outward side, tau_w and q_w
   (n_t1, n_t2).  w (5, N...) ghosted primitives and T (N...).  The block descriptor
   supplies n, the face's T_wall / n_match and the geometry (S, xf, xc).  Replaces
   WallModel.wall_flux(w, T, blk, geom, gas, axis, side) (reference wmles/bc.py:145-166). */
int wmx_wall_flux(const wmx_block_desc* block, int32_t axis, int32_t side, const wmx_numerics_desc* numerics,
                  const double* w, const double* T, double* F, double* tau_w, double* q_w, void* stream);

/* ---- learned wall model (no counterpart in the reference; see wmx_wall_net_desc) ---- */
/* Parameter counts: FP32 master, BF16 inference copy, FP32 inference copy. */
int wmx_mlp_sizes(int32_t* nparam, int32_t* nbf16, int32_t* nfp32);
/* n training samples of the method: features uniform over net->lo..hi, u+ by the
   Newton of wmles/wall_model.py:19-40, T+ by Kader (wall_model.py:56-60).  X (n, 3)
   and Y (n, 2) normalised with net's centers/scales; raw (n, 5) optional:
   log Re_h, h/Delta, T_h/T_w, u+, T+.  Deterministic in (seed, index). */
int wmx_wm_dataset(const wmx_numerics_desc* numerics, const wmx_wall_net_desc* net, int64_t n,
                   uint64_t seed, float* X, float* Y, double* raw, void* stream);
/* One FP32 forward + backward over samples [offset, offset + batch) of (X, Y)
   (mean squared error over batch x 2 outputs).  partials: ceil(batch/64) * nparam
   floats; loss_partials: ceil(batch/64) sums of squared errors.  grad (optional):
   the summed gradient.  update != 0: Adam step `step` (>= 1) with rate lr on params. */
int wmx_mlp_train_step(float* params, const float* X, const float* Y, int64_t n_samples,
                       int64_t offset, int32_t batch, float* partials, float* loss_partials,
                       float* grad, float* adam_m, float* adam_v, int32_t step, float lr,
                       int32_t update, void* stream);
/* FP32 master -> inference copies (BF16 W2 | W3, FP32 W1 | b1 | b2 | b3). */
int wmx_mlp_pack(const float* params, uint16_t* w_bf16, float* w_fp32, void* stream);
/* FP32 forward of the master parameters on n normalised inputs -> Y (n, 2). */
int wmx_mlp_forward(const float* params, const float* X, int64_t n, float* Y, void* stream);
/* The learned model on n points: the wall-model callable of wmles/wall_model.py:42-61
   with the Newton start taken from the BF16 MFMA network.  mode 0: n_polish Newton
   steps, residual check, Newton fallback, q_w by Kader at the polished y+; mode 1:
   network outputs only (u+ and T+).  hdel optional (h/Delta).  uplus optional. */
int wmx_wall_net_eval(const wmx_numerics_desc* numerics, const wmx_wall_net_desc* net,
                      int32_t mode, int64_t n, const double* h, const double* hdel,
                      const double* U, const double* T_h, const double* p_h, double T_wall,
                      double* tau_w, double* q_w, double* uplus, void* stream);
/* Use the learned model (mode 0) in the plan's wall-flux kernel; net = NULL restores
   the Newton model.  The descriptor's device buffers must outlive the plan's use. */
int wmx_plan_set_wall_net(wmx_plan* plan, const wmx_wall_net_desc* net);
/* D = A B for one 16x16x16 tile through v_mfma_f32_16x16x16bf16_1k with the lane
   layout the network kernels use (A, B row-major FP32, rounded to BF16; D FP32). */
int wmx_mfma_bf16_probe(const float* A, const float* B, float* D, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* WMLES_AMD_H */

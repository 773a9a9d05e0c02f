// C ABI of libwmles_amd.so: plan construction (ghost-sweep phases, tile tables,
// workspace layout) and the launch sequences of the explicit step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

using namespace wmx;

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return 1;
}

#define WMX_CHECK(expr)                                                            \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr size_t ALIGN = 256;

size_t align_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

struct Box {
  int blk;
  int lo[3], hi[3];
};

bool overlap(const Box& a, const Box& b) {
  if (a.blk != b.blk) return false;
  for (int d = 0; d < 3; ++d)
    if (a.hi[d] <= b.lo[d] || b.hi[d] <= a.lo[d]) return false;
  return true;
}

Box full_box(int blk, const int* N) {
  Box b{blk, {0, 0, 0}, {N[0], N[1], N[2]}};
  return b;
}

template <class T>
int upload(const std::vector<T>& v, T** d) {
  *d = nullptr;
  if (v.empty()) return 0;
  WMX_CHECK(hipMalloc((void**)d, sizeof(T) * v.size()));
  WMX_CHECK(hipMemcpy(*d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

struct wmx_plan {
  int nb = 0, ng = 3;
  std::vector<wmx_block_desc> desc;
  std::vector<BlockDev> hb;
  BlockDev* d_blks = nullptr;
  NumDev nd{};
  int n_cases = 1;
  bool any_constraint = false;
  std::vector<int> case_dir;
  std::vector<double> case_target;
  // ghost sweep
  int passes = 0;                     // sweep passes kept after the redundancy check
  std::vector<GhostOp> ops;           // all phases, flattened
  std::vector<int> phase_off;         // size nphases + 1
  std::vector<int> phase_max;
  GhostOp* d_ops = nullptr;
  std::vector<std::vector<int>> phase_inflow;  // block ids of inflow fills per phase
  std::vector<GhostOp> phase_ext;     // per phase: the external fill it stands for (blk -1: none)
  bool has_external = false;          // some face is WMX_BC_EXTERNAL: the caller runs the sweep
  bool fuse_last = false;   // the last phase is k-face fills only: a step's primitive launches
                            // evaluate those ghosts (BlockDev::zfuse) and skip its fill launch
  // sponge row means in head workgroups of the primitive launch (every block's rows of
  // 9..129 cells): per-block first workgroup, 32 rows each
  std::vector<int> sponge_pre;
  int* d_sponge_pre = nullptr;
  // wall faces
  std::vector<WallOp> wall;
  WallOp* d_wall = nullptr;
  int wall_max = 0;
  // flux tiles
  std::vector<int4> tiles;
  int4* d_tiles = nullptr;
  int ci_req = 0;                     // requested i-planes per marching tile (0: auto)
  int ci_len = 16;                    // i-planes per marching tile in use
  int tile_order = 1;                 // 0: block/i/j/k order, 1: XCD-interleaved
  int cu_count = 256;                 // compute units of the device the plan was made on
  // captured steps: (q0, qout, d_dt, time) -> graph of max_dt + advance
  struct GraphEntry {
    std::vector<double*> q0, qout;
    const double* d_dt;
    double time;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  bool capturing = false;
  // per-kernel workgroup prefix tables
  std::vector<int> pre_all, pre_box, pre_int, pre_shell;
  int *d_pre_all = nullptr, *d_pre_box = nullptr, *d_pre_int = nullptr, *d_pre_shell = nullptr;

  // case tables over the interior-cell workgroups
  std::vector<int> case_wg, case_off;
  int *d_case_wg = nullptr, *d_case_off = nullptr, *d_case_dir = nullptr;
  // the same per-case workgroup lists for the all-cells layout (fused stage-1 prim + dt)
  std::vector<int> case_wg_all, case_off_all;
  int *d_case_wg_all = nullptr, *d_case_off_all = nullptr;
  // a list that is the identity (blocks ordered by case) is passed as nullptr
  bool case_wg_ident = false, case_wg_all_ident = false;
  double* d_case_target = nullptr;
  // inflows
  std::vector<InflowDev> inflows;
  // workspace
  size_t ws_bytes = 0;
  std::vector<size_t> off_P, off_G, off_shell, off_sponge, off_geo32, off_sb;
  std::vector<std::vector<long long>> off_wall;   // per block, 6 faces (-1: none)
  size_t off_part = 0, off_scal = 0, off_stat = 0;
  double* d_stat = nullptr;           // plane-statistics partials
  char* ws = nullptr;
  double* d_part = nullptr;
  double* d_delta = nullptr;
  // persistent stage launches (FP32-bulk plans without inflow, sponge, learned wall model
  // or caller-filled faces): one cooperative launch per RK stage
  bool fuse_grad = true;              // FP32-bulk: gradient records evaluated in the flux march
  bool persistent = false;
  bool stage_dirty = true;            // d_stage needs re-uploading (tiling changed)
  StageStatic* d_stage = nullptr;
  Ptrs* d_pbuf = nullptr;
  int* d_fill_off = nullptr;
  int* d_fill_max = nullptr;
  int stage_grid = 0, stage_grid_f = 0;   // persistent grids in use (0: not yet launched)
  // kernel-class event timing
  int prof_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_rec;   // (class, index of start event)
};

namespace {

int nwg_of(const std::vector<int>& pre) { return pre.empty() ? 0 : pre.back(); }

enum { K_FLUX = 0, K_GRAD, K_PRIM, K_FILL, K_WALL, K_DT, K_MF, K_SHELL, K_INFLOW, K_SPONGE, K_NCLASS };

hipEvent_t next_event(wmx_plan* p) {
  if (p->ev_used == p->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p->ev_pool.push_back(e);
  }
  return p->ev_pool[p->ev_used++];
}

// bracket one launch of class `cls` with events when profiling is on
template <class F>
hipError_t timed(wmx_plan* p, int cls, hipStream_t s, F&& launch) {
  if (!(p->prof_mask & (1 << cls)) || p->capturing) return launch();
  const size_t i0 = p->ev_used;
  hipEvent_t a = next_event(p), b = next_event(p);
  if (!a || !b) return launch();
  hipEventRecord(a, s);
  hipError_t e = launch();
  hipEventRecord(b, s);
  p->ev_rec.push_back({cls, i0});
  return e;
}

std::vector<int> prefix(const std::vector<long long>& counts) {
  std::vector<int> p(counts.size() + 1, 0);
  for (size_t i = 0; i < counts.size(); ++i) p[i + 1] = p[i] + (int)((counts[i] + 255) / 256);
  return p;
}

// read / write boxes of one fill op (sequential-sweep semantics of solver.py:59-73)
void op_boxes(const wmx_plan& p, int b, int a, int s, std::vector<Box>* rd, Box* wr) {
  const BlockDev& B = p.hb[b];
  const int ng = B.ng, n = B.n[a];
  *wr = full_box(b, B.N);
  wr->lo[a] = s == 0 ? 0 : ng + n;
  wr->hi[a] = s == 0 ? ng : B.N[a];
  const wmx_face_desc& f = B.faces[2 * a + s];
  rd->clear();
  if (f.kind == WMX_FACE_INTERFACE) {
    const BlockDev& S = p.hb[f.nb];
    Box r = full_box(f.nb, S.N);
    const int aB = f.src_axis, nB = S.n[aB];
    r.lo[aB] = f.src_side == 0 ? ng : nB;
    r.hi[aB] = r.lo[aB] + ng;
    rd->push_back(r);
  } else if (f.kind == WMX_BC_DIRICHLET || f.kind == WMX_BC_INFLOW) {
    // reads no state
  } else if (f.kind == WMX_BC_EXTERNAL) {
    rd->push_back(full_box(b, B.N));      // a caller's fill may read anything in its block
  } else {
    Box r = full_box(b, B.N);
    if (f.kind == WMX_BC_EXTRAPOLATE) {
      r.lo[a] = s == 0 ? ng : ng + n - 1;
      r.hi[a] = r.lo[a] + 1;
    } else {
      r.lo[a] = s == 0 ? ng : n;
      r.hi[a] = r.lo[a] + ng;
    }
    rd->push_back(r);
  }
}

// ---- redundant sweep passes ------------------------------------------------------
// Executes the sweep symbolically on 64-bit value signatures: interior cells and the
// initial ghost values get unique ids; an interface copy carries its source's
// signature; a boundary fill hashes (op, destination layer position, source
// signature).  Equal signatures mean equal expression trees, hence bit-identical
// floating-point values.  Passes after the first one that leaves every signature
// unchanged are dropped: they would rewrite the same values.
inline uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

int effective_passes(const wmx_plan& p, int passes) {
  if (passes <= 1) return passes;
  const int nb = p.nb;
  std::vector<std::vector<uint64_t>> sig(nb);
  uint64_t next = 1;
  for (int b = 0; b < nb; ++b) {
    sig[b].resize((size_t)p.hb[b].cells);
    for (auto& v : sig[b]) v = mix64(next++);
  }
  auto run_pass = [&]() {
    int opi = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < nb; ++b)
        for (int s = 0; s < 2; ++s, ++opi) {
          const BlockDev& B = p.hb[b];
          const wmx_face_desc& f = B.faces[2 * a + s];
          const int ng = B.ng, n = B.n[a];
          int t1, t2;
          tangents(a, &t1, &t2);
          for (int m = 0; m < ng; ++m)
            for (int i1 = 0; i1 < B.N[t1]; ++i1)
              for (int i2 = 0; i2 < B.N[t2]; ++i2) {
                int c[3];
                c[t1] = i1; c[t2] = i2;
                c[a] = s == 1 ? ng + n + m : ng - 1 - m;
                const size_t di = (size_t)cidx(B.N, c[0], c[1], c[2]);
                uint64_t v;
                if (f.kind == WMX_FACE_INTERFACE) {
                  const BlockDev& S = p.hb[f.nb];
                  int cs[3];
                  for (int t = 0; t < 3; ++t) {
                    if (t == a) continue;
                    cs[f.mapping[t]] = f.flip[t] ? B.N[t] - 1 - c[t] : c[t];
                  }
                  const int nB = S.n[f.src_axis];
                  cs[f.src_axis] = f.src_side == 1 ? nB + ng - 1 - m : ng + m;
                  v = sig[f.nb][(size_t)cidx(S.N, cs[0], cs[1], cs[2])];
                } else {
                  int ci[3] = {c[0], c[1], c[2]};
                  ci[a] = s == 0 ? ng + m : ng + n - 1 - m;
                  if (f.kind == WMX_BC_EXTRAPOLATE) ci[a] = s == 0 ? ng : ng + n - 1;
                  uint64_t src = 0;
                  if (f.kind != WMX_BC_DIRICHLET && f.kind != WMX_BC_INFLOW)
                    src = sig[b][(size_t)cidx(B.N, ci[0], ci[1], ci[2])];
                  v = mix64(mix64(((uint64_t)opi << 40) ^ di) ^ src);
                }
                sig[b][di] = v;
              }
        }
  };
  run_pass();
  for (int k = 1; k < passes; ++k) {
    std::vector<std::vector<uint64_t>> prev = sig;
    run_pass();
    if (prev == sig) return k;          // pass k+1 reproduced pass k exactly
  }
  return passes;
}

int build_phases(wmx_plan& p, int passes) {
  struct Pending {
    GhostOp op;
    std::vector<Box> rd;
    Box wr;
  };
  std::vector<Pending> cur;
  p.ops.clear();
  p.phase_off.assign(1, 0);
  p.phase_max.clear();
  p.phase_inflow.clear();
  p.phase_ext.clear();
  auto close = [&]() {
    if (cur.empty()) return;
    p.phase_ext.push_back(GhostOp{-1, -1, -1, 0});
    int mx = 0;
    std::vector<int> inflow;
    for (auto& c : cur) {
      p.ops.push_back(c.op);
      mx = std::max(mx, c.op.count);
      if (p.hb[c.op.blk].faces[2 * c.op.axis + c.op.side].kind == WMX_BC_INFLOW)
        inflow.push_back(c.op.blk);
    }
    p.phase_off.push_back((int)p.ops.size());
    p.phase_max.push_back(mx);
    p.phase_inflow.push_back(inflow);
    cur.clear();
  };
  for (int pass = 0; pass < passes; ++pass)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < p.nb; ++b)
        for (int s = 0; s < 2; ++s) {
          Pending x;
          const BlockDev& B = p.hb[b];
          int t1, t2;
          tangents(a, &t1, &t2);
          x.op = GhostOp{b, a, s, B.ng * B.N[t1] * B.N[t2]};
          if (B.faces[2 * a + s].kind == WMX_BC_EXTERNAL) {
            // a phase of its own, with no op for the fill kernel
            close();
            p.phase_off.push_back((int)p.ops.size());
            p.phase_max.push_back(0);
            p.phase_inflow.push_back({});
            p.phase_ext.push_back(x.op);
            continue;
          }
          op_boxes(p, b, a, s, &x.rd, &x.wr);
          bool conflict = false;
          for (auto& c : cur) {
            if (overlap(c.wr, x.wr)) conflict = true;
            for (auto& r : x.rd) if (overlap(c.wr, r)) conflict = true;
            for (auto& r : c.rd) if (overlap(x.wr, r)) conflict = true;
            if (conflict) break;
          }
          if (conflict) close();
          cur.push_back(x);
        }
  close();
  return 0;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Flux tiles: per block (i-chunk, j-tile, k-tile).  Order 1 interleaves contiguous
// spans of the tile list across the 8 XCDs (workgroup w runs on XCD w % 8), so the
// tiles one XCD runs are neighbours and share its L2.
// i-planes per tile when none is requested: both flux kernels run two 256-thread
// workgroups per CU, so a launch of T tiles takes ceil(T / (2 CUs)) rounds of
// (ci + alpha) plane times, alpha being the per-tile start-up (the FP32-bulk kernel's
// register window costs about 1.5 planes, the FP64 kernel's about 0.5).  The shortest
// estimate over ci in [4, 64] wins (ties: the longer chunk).  dims: n[3] per block.
int auto_ci_len(int nb, const int* dims, int prec, int cu_count, long long* ntiles) {
  const int tjr = FLUX2_TILE_J;
  const double alpha = prec == WMX_PREC_FP32_BULK ? 1.5 : 0.5;
  const long long slots = 2LL * std::max(1, cu_count);
  int best = 16;
  long long best_T = 0;
  double best_cost = 1e300;
  for (int ci = 64; ci >= 4; --ci) {
    long long T = 0;
    for (int b = 0; b < nb; ++b) {
      const int* n = dims + 3 * b;
      T += (long long)((n[0] + ci - 1) / ci) * ((n[1] + tjr - 1) / tjr) * ((n[2] + 15) / 16);
    }
    const double cost = (double)((T + slots - 1) / slots) * (ci + alpha);
    if (cost < best_cost - 1e-9) { best_cost = cost; best = ci; best_T = T; }
  }
  if (ntiles) *ntiles = best_T;
  return best;
}

void build_tiles(wmx_plan* p) {
  std::vector<int4> t;
  // tile rows of both flux kernels
  const int tjr = FLUX2_TILE_J;
  if (p->ci_req > 0) {
    p->ci_len = p->ci_req;
  } else {
    std::vector<int> dims;
    for (int b = 0; b < p->nb; ++b) dims.insert(dims.end(), p->hb[b].n, p->hb[b].n + 3);
    p->ci_len = auto_ci_len(p->nb, dims.data(), p->nd.prec, p->cu_count, nullptr);
  }
  for (int b = 0; b < p->nb; ++b) {
    const BlockDev& B = p->hb[b];
    for (int i0 = 0; i0 < B.n[0]; i0 += p->ci_len)
      for (int j0 = 0; j0 < B.n[1]; j0 += tjr)
        for (int k0 = 0; k0 < B.n[2]; k0 += 16) t.push_back(make_int4(b, j0, k0, i0));
  }
  if (p->tile_order == 1) {
    const size_t n = t.size(), nx = 8, span = (n + nx - 1) / nx;
    std::vector<int4> o;
    o.reserve(n);
    for (size_t r = 0; r < span; ++r)
      for (size_t x = 0; x < nx; ++x) {
        const size_t src = x * span + r;
        if (src < n) o.push_back(t[src]);
      }
    t.swap(o);
  }
  p->tiles.swap(t);
}

NumDev make_numdev(const wmx_numerics_desc* nm) {
  NumDev nd;
  std::memset(&nd, 0, sizeof(nd));
  nd.gas = GasC{nm->gamma, nm->R, nm->Pr, nm->Pr_t, nm->mu_ref, nm->T_ref, nm->S_suth,
                nm->gamma * nm->R / (nm->gamma - 1.0)};
  nd.rc = RecC{1.0 / nm->R, 1.0 / (nm->gamma - 1.0), (float)(1.0 / nm->T_ref),
               (float)(nm->mu_ref * (nm->T_ref + nm->S_suth)), (float)nm->S_suth};
  nd.wm = WallModelC{nm->kappa, nm->B, nm->wm_ekb, nm->wm_pr13, nm->wm_pr3, nm->wm_beta, nm->n_iter};
  nd.scheme = nm->scheme;
  nd.central = nm->central;
  nd.order = nm->order;
  nd.limiter = nm->limiter;
  nd.th0 = nm->th0;
  nd.th1 = nm->th1;
  nd.sgs = nm->sgs;
  nd.sgs_c = nm->sgs_c;
  nd.cfl = nm->cfl;
  nd.n_sponge = nm->n_sponge;
  std::memcpy(nd.sponge, nm->sponge, sizeof(nd.sponge));
  nd.prec = nm->precision;
  return nd;
}

Ptrs make_ptrs(const wmx_plan* p, double* const* q) {
  Ptrs r;
  std::memset(&r, 0, sizeof(r));
  for (int b = 0; b < p->nb; ++b) r.p[b] = q[b];
  return r;
}

// zfuse: the last phase is left to the primitive launch (plans with fuse_last)
int run_fill(wmx_plan* p, Ptrs q, double time, const double* d_dt, double coef, hipStream_t s,
             bool zfuse = false) {
  const int nph = (int)p->phase_max.size() - (zfuse && p->fuse_last ? 1 : 0);
  for (int ph = 0; ph < nph; ++ph) {
    if (p->phase_ext[ph].blk >= 0) continue;          // the caller's fill
    const int o = p->phase_off[ph], n = p->phase_off[ph + 1] - o;
    WMX_CHECK(timed(p, K_FILL, s, [&] { return launch_fill_cells(p->d_blks, p->d_ops + o, n, p->phase_max[ph], q, p->nd.gas, s); }));
    for (int blk : p->phase_inflow[ph]) {
      const BlockDev& B = p->hb[blk];
      const int id = B.faces[0].inflow_id;
      WMX_CHECK(timed(p, K_INFLOW, s, [&] {
        return launch_inflow(p->d_blks, blk, B.N[1], B.N[2], B.ng, p->inflows[id], q, p->nd.gas, time,
                             d_dt, coef, s);
      }));
    }
  }
  return 0;
}

// ghost fill + cell-level fields of one RHS evaluation
// dt_out != nullptr: the primitive launch also takes the CFL partials of the (unchanged)
// interior state and the per-case minimum goes to dt_out before the gradients
// zfuse: a step's stage (the sweep's last phase evaluated by the primitive launch)
int run_cells(wmx_plan* p, Ptrs q, double time, const double* d_dt, double coef, hipStream_t s,
              double* dt_out = nullptr, bool zfuse = false, bool fill = true) {
  zfuse = zfuse && p->fuse_last && fill;
  if (fill && run_fill(p, q, time, d_dt, coef, s, zfuse)) return 1;
  SpongeHead sh;
  if (p->d_sponge_pre) {
    sh.pre = p->d_sponge_pre;
    sh.nwg = p->sponge_pre.back();
  }
  // fused FP32-bulk plans: the Newton wall faces ride in the primitive launch's head
  const bool f32 = p->nd.prec == WMX_PREC_FP32_BULK;
  const bool wall_in_prim = f32 && p->fuse_grad && !p->nd.net.w && !p->wall.empty() && p->wall_max > 0;
  if (wall_in_prim) {
    sh.wall = p->d_wall;
    sh.wall_wgx = (p->wall_max + 255) / 256;
    sh.nwall = (int)p->wall.size();
  }
  WMX_CHECK(timed(p, K_PRIM, s, [&] {
    return launch_prim(p->d_blks, p->d_pre_all, p->nb, nwg_of(p->pre_all), q, p->nd,
                       dt_out ? p->d_part : nullptr, zfuse ? 1 : 0, sh, s);
  }));
  // FP32-bulk plans take the per-case dt mins and (Newton model, unfused) the wall faces into
  // the gradient launch; the flux kernel after it is the first reader of both
  const bool wall_in_grad = f32 && !p->nd.net.w && !wall_in_prim;
  GradDt gdt;
  if (dt_out && f32) {
    gdt.ncases = p->n_cases;
    gdt.list = p->case_wg_all_ident ? nullptr : p->d_case_wg_all;
    gdt.off = p->d_case_off_all;
    gdt.part = p->d_part;
    gdt.out = dt_out;
  } else if (dt_out) {
    WMX_CHECK(timed(p, K_DT, s, [&] {
      return launch_case_final(0, p->n_cases, p->case_wg_all_ident ? nullptr : p->d_case_wg_all,
                               p->d_case_off_all, p->d_part, nullptr, dt_out, s);
    }));
  }
  // fused FP32-bulk plans: the flux kernel evaluates the gradient records itself; the
  // gradient launch keeps only its head work (per-case dt min, Newton wall faces)
  const bool grads = !(f32 && p->fuse_grad);
  WMX_CHECK(timed(p, K_GRAD, s, [&] {
    return launch_grad(p->d_blks, p->d_pre_box, p->nb, grads ? nwg_of(p->pre_box) : 0, p->nd,
                       wall_in_grad ? p->d_wall : nullptr, (int)p->wall.size(), p->wall_max, q, gdt, s);
  }));
  if (!wall_in_grad && !wall_in_prim)
    WMX_CHECK(timed(p, K_WALL, s, [&] { return launch_wall(p->d_blks, p->d_wall, (int)p->wall.size(), p->wall_max, q, p->nd, s); }));
  if (p->nd.n_sponge > 0 && !p->d_sponge_pre)
    for (int b = 0; b < p->nb; ++b)
      WMX_CHECK(timed(p, K_SPONGE, s, [&] {
        return launch_sponge_mean(p->d_blks, b, (long long)p->hb[b].n[0] * p->hb[b].n[1], q, s);
      }));
  return 0;
}

hipError_t plan_flux(const wmx_plan* p, int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles,
                     int ci_len, const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt, double dt,
                     hipStream_t s) {
  if (nd.prec == WMX_PREC_FP32_BULK && p->fuse_grad)
    return launch_flux3(mode, d_blks, d_tiles, ntiles, ci_len, nd, qin, q0, out, d_dt, dt, s);
  return launch_flux(mode, d_blks, d_tiles, ntiles, ci_len, nd, qin, q0, out, d_dt, dt, s);
}

int check_plan(const wmx_plan* p) {
  if (!p) return fail("null plan");
  if (!p->ws) return fail("plan has no workspace bound (wmx_plan_bind_workspace)");
  return 0;
}

}  // namespace

void drop_graphs(wmx_plan* p) {
  for (auto& g : p->graphs) (void)hipGraphExecDestroy(g.exec);
  p->graphs.clear();
}

int check_net(const wmx_wall_net_desc* n) {
  if (!n->w_bf16 || !n->w_fp32) return fail("wall net: null weight pointer");
  for (int k = 0; k < 3; ++k)
    if (!(n->hi[k] > n->lo[k]) || !(n->x_scale[k] > 0.f)) return fail("wall net: empty input range");
  for (int k = 0; k < 2; ++k)
    if (!(n->y_scale[k] > 0.f)) return fail("wall net: y_scale must be > 0");
  if (n->n_polish < 0 || n->n_polish > 100) return fail("wall net: n_polish must be in [0, 100]");
  if (!(n->tol > 0.0)) return fail("wall net: tol must be > 0");
  return 0;
}

WallNetDev make_netdev(const wmx_wall_net_desc* n) {
  WallNetDev d;
  std::memset(&d, 0, sizeof(d));
  d.w = n->w_bf16;
  d.f = n->w_fp32;
  for (int k = 0; k < 3; ++k) {
    d.xc[k] = n->x_center[k];
    d.xs[k] = n->x_scale[k];
    d.lo[k] = n->lo[k];
    d.hi[k] = n->hi[k];
  }
  for (int k = 0; k < 2; ++k) {
    d.yc[k] = n->y_center[k];
    d.ys[k] = n->y_scale[k];
  }
  d.n_polish = n->n_polish;
  d.tol = n->tol;
  d.nfb = n->fallback_count;
  return d;
}

extern "C" {

const char* wmx_last_error(void) { return g_err.c_str(); }

const char* wmx_build_info(void) {
  return "libwmles_amd: gfx950, FP64 plain form, hipcc -O3, no packed-fp32";
}

int wmx_plan_create(int32_t nblocks, const wmx_block_desc* blocks, const wmx_numerics_desc* nm,
                    int32_t ninflow, const wmx_inflow_desc* inflows, wmx_plan** out) {
  if (!out || !blocks || !nm) return fail("null argument");
  *out = nullptr;
  if (nblocks < 1 || nblocks > WMX_MAX_BLOCKS)
    return fail("nblocks must be in [1, " + std::to_string(WMX_MAX_BLOCKS) + "]");
  if (nm->n_cases < 1 || nm->n_cases > 16) return fail("n_cases must be in [1, 16]");
  if (nm->n_sponge < 0 || nm->n_sponge > 4) return fail("at most 4 sponge sources");
  if (nm->order != 2 && nm->order != 4) return fail("central order must be 2 or 4");
  if (nm->bc_passes < 1) return fail("bc_passes must be >= 1");
  if (nm->precision != WMX_PREC_FP64 && nm->precision != WMX_PREC_FP32_BULK)
    return fail("precision must be WMX_PREC_FP64 or WMX_PREC_FP32_BULK");
  wmx_plan* p = new wmx_plan();
  p->nb = nblocks;
  p->desc.assign(blocks, blocks + nblocks);
  NumDev& nd = p->nd;
  nd = make_numdev(nm);
  p->n_cases = nm->n_cases;
  p->case_dir.assign(nm->mass_flux_dir, nm->mass_flux_dir + 16);
  p->case_target.assign(nm->mass_flux_target, nm->mass_flux_target + 16);
  for (int c = 0; c < p->n_cases; ++c) p->any_constraint |= p->case_dir[c] >= 0;
  // blocks
  p->hb.resize(nblocks);
  const int ng = 3;
  p->ng = ng;
  for (int b = 0; b < nblocks; ++b) {
    const wmx_block_desc& d = blocks[b];
    BlockDev& B = p->hb[b];
    std::memset(&B, 0, sizeof(B));
    for (int a = 0; a < 3; ++a) {
      if (d.n[a] < ng) { delete p; return fail("each block needs at least ng cells per direction"); }
      B.n[a] = d.n[a];
      B.N[a] = d.n[a] + 2 * ng;
    }
    B.ng = ng;
    B.case_id = d.case_id;
    if (d.case_id < 0 || d.case_id >= p->n_cases) { delete p; return fail("case_id out of range"); }
    B.cells = (long long)B.N[0] * B.N[1] * B.N[2];
    // kernels address each plane / face-component array with 32-bit byte offsets
    if ((long long)(B.N[0] + 1) * (B.N[1] + 1) * (B.N[2] + 1) * (long long)sizeof(double) >= (1LL << 32)) {
      delete p;
      return fail("block too large: every FP64 plane (cells or faces incl. ghosts) must stay below 4 GiB");
    }
    B.vol = d.vol;
    B.xc = d.xc;
    for (int a = 0; a < 3; ++a) { B.S[a] = d.S[a]; B.xf[a] = d.xf[a]; }
    if (!B.vol || !B.xc || !B.S[0] || !B.S[1] || !B.S[2] || !B.xf[0] || !B.xf[1] || !B.xf[2]) {
      delete p;
      return fail("block geometry pointers must be set");
    }
    B.shell_cells = shell_count(B.N, B.n, ng);
    for (int f = 0; f < 6; ++f) B.faces[f] = d.faces[f];
  }
  // validate faces
  for (int b = 0; b < nblocks; ++b)
    for (int f = 0; f < 6; ++f) {
      const wmx_face_desc& fc = p->hb[b].faces[f];
      const int a = f / 2;
      if (fc.kind == WMX_FACE_INTERFACE) {
        if (fc.nb < 0 || fc.nb >= nblocks) { delete p; return fail("interface source block out of range"); }
        const BlockDev& S = p->hb[fc.nb];
        for (int t = 0; t < 3; ++t) {
          if (t == a) continue;
          if (S.N[fc.mapping[t]] != p->hb[b].N[t]) { delete p; return fail("interface extents do not match"); }
        }
        if (fc.mapping[a] != fc.src_axis) { delete p; return fail("interface mapping inconsistent"); }
      } else if (fc.kind == WMX_BC_INFLOW) {
        if (f != 0) { delete p; return fail("inflow is only valid on the i-min face"); }
        if (fc.inflow_id < 0 || fc.inflow_id >= ninflow) { delete p; return fail("inflow_id out of range"); }
        if (p->hb[b].N[2] > 1024) { delete p; return fail("inflow fill supports N_k <= 1024"); }
      } else if (fc.kind == WMX_BC_WALLMODEL) {
        if (fc.n_match < 1 || fc.n_match > p->hb[b].n[a]) { delete p; return fail("n_match out of range"); }
      } else if (fc.kind == WMX_BC_EXTERNAL) {
        p->has_external = true;
      } else if (fc.kind < 0 || fc.kind > WMX_BC_EXTERNAL) {
        delete p;
        return fail("unknown face kind");
      }
    }
  for (int i = 0; i < ninflow; ++i) {
    const wmx_inflow_desc& in = inflows[i];
    if (in.n_tab < 1 || in.n_modes < 1 || !in.table || !in.modes) {
      delete p;
      return fail("inflow needs a profile table and at least one mode");
    }
    p->inflows.push_back(InflowDev{in.n_tab, in.n_modes, 0, in.table, in.modes, in.U_inf, in.L,
                                   in.rho_inf, in.p_inf});
  }
  // (what a caller's fill writes is unknown: every pass of such a plan is kept)
  p->passes = p->has_external ? nm->bc_passes : effective_passes(*p, nm->bc_passes);
  build_phases(*p, p->passes);
  // The last phase's values are read by nothing else in a step but the primitive kernel
  // (the later kernels read the primitive records or interior states; the step's output
  // ghosts come from the saved shell), so when it holds only k-face, non-inflow fills the
  // primitive kernels of the step's stages evaluate them in place of its launch.
  {
    const int nph = (int)p->phase_max.size();
    bool ok = nph > 0 && !p->has_external;
    if (ok)
      for (int o = p->phase_off[nph - 1]; o < p->phase_off[nph]; ++o) {
        const GhostOp& op = p->ops[o];
        if (op.axis != 2 || p->hb[op.blk].faces[4 + op.side].kind == WMX_BC_INFLOW) ok = false;
        // no earlier phase writes those k ghosts either (then a step's first stage leaves
        // them unwritten in its input, which the ghost-shell kernels rely on)
        for (int e = 0; ok && e < p->phase_off[nph - 1]; ++e)
          if (p->ops[e].axis == 2 && p->ops[e].blk == op.blk && p->ops[e].side == op.side) ok = false;
      }
    p->fuse_last = ok;
    if (ok)
      for (int o = p->phase_off[nph - 1]; o < p->phase_off[nph]; ++o)
        p->hb[p->ops[o].blk].zfuse |= 1 << p->ops[o].side;
  }
  // wall faces
  for (int b = 0; b < nblocks; ++b)
    for (int f = 0; f < 6; ++f)
      if (p->hb[b].faces[f].kind == WMX_BC_WALLMODEL) {
        int t1, t2;
        tangents(f / 2, &t1, &t2);
        int cnt = p->hb[b].n[t1] * p->hb[b].n[t2];
        p->wall.push_back(WallOp{b, f / 2, f % 2, cnt});
        p->wall_max = std::max(p->wall_max, cnt);
      }
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      p->cu_count = cus;
  }
  build_tiles(p);
  // prefix tables
  std::vector<long long> c_all, c_box, c_int, c_shell;
  for (auto& B : p->hb) {
    c_all.push_back(B.cells);
    const int run = grad_run(nd.prec == WMX_PREC_FP32_BULK ? sizeof(float) : sizeof(double));
    c_box.push_back((long long)((B.n[0] + 2 + run - 1) / run) * (B.n[1] + 2) * (B.n[2] + 2));
    c_int.push_back(((long long)B.n[0] * B.n[1] * B.n[2] + INT_CELLS - 1) / INT_CELLS);   // thread slots
    c_shell.push_back(B.shell_cells);
  }
  p->pre_all = prefix(c_all);
  p->pre_box = prefix(c_box);
  p->pre_int = prefix(c_int);
  p->pre_shell = prefix(c_shell);
  p->case_off.assign(1, 0);
  for (int c = 0; c < p->n_cases; ++c) {
    for (int b = 0; b < nblocks; ++b)
      if (p->hb[b].case_id == c)
        for (int w = p->pre_int[b]; w < p->pre_int[b + 1]; ++w) p->case_wg.push_back(w);
    p->case_off.push_back((int)p->case_wg.size());
  }
  p->case_off_all.assign(1, 0);
  for (int c = 0; c < p->n_cases; ++c) {
    for (int b = 0; b < nblocks; ++b)
      if (p->hb[b].case_id == c)
        for (int w = p->pre_all[b]; w < p->pre_all[b + 1]; ++w) p->case_wg_all.push_back(w);
    p->case_off_all.push_back((int)p->case_wg_all.size());
  }
  auto ident = [](const std::vector<int>& v) {
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i] != (int)i) return false;
    return true;
  };
  p->case_wg_ident = ident(p->case_wg);
  p->case_wg_all_ident = ident(p->case_wg_all);
  // sponge means in the primitive launch when every block's k rows suit the 8-lane form
  if (nd.n_sponge > 0) {
    bool ok = true;
    for (const BlockDev& B : p->hb) ok = ok && B.n[2] - 1 >= 8 && B.n[2] - 1 <= 128;
    if (ok) {
      p->sponge_pre.assign(1, 0);
      for (const BlockDev& B : p->hb)
        p->sponge_pre.push_back(p->sponge_pre.back() + (int)(((long long)B.n[0] * B.n[1] + 31) / 32));
    }
  }
  // workspace layout
  size_t off = 0;
  p->off_wall.assign(nblocks, std::vector<long long>(6, -1));
  for (int b = 0; b < nblocks; ++b) {
    const BlockDev& B = p->hb[b];
    const size_t rec = nd.prec == WMX_PREC_FP32_BULK ? sizeof(float) : sizeof(double);
    p->off_P.push_back(off);
    off = align_up(off + rec * NP * B.cells);
    p->off_G.push_back(off);
    off = align_up(off + rec * (nd.prec == WMX_PREC_FP32_BULK ? 4 * G32_GROUPS : NGR) * B.cells);
    p->off_geo32.push_back(off);
    if (nd.prec == WMX_PREC_FP32_BULK)
      off = align_up(off + sizeof(float) * (3 * (size_t)(fcount(B.N, 0) + fcount(B.N, 1) + fcount(B.N, 2)) +
                                            (size_t)B.cells));
    p->off_sb.push_back(off);
    off = align_up(off + sizeof(double) * NSB_PLANES * (size_t)B.cells);
    p->off_shell.push_back(off);
    off = align_up(off + sizeof(double) * 5 * B.shell_cells);
    p->off_sponge.push_back(off);
    if (nd.n_sponge > 0)
      off = align_up(off + (sizeof(double) * 5 + sizeof(int)) * (size_t)B.n[0] * B.n[1]);
    for (int f = 0; f < 6; ++f)
      if (B.faces[f].kind == WMX_BC_WALLMODEL) {
        int t1, t2;
        tangents(f / 2, &t1, &t2);
        p->off_wall[b][f] = (long long)off;
        off = align_up(off + sizeof(double) * WALL_PLANES * B.n[t1] * B.n[t2]);
      }
  }
  p->off_part = off;
  // partials: two per interior workgroup (mass flux), one per all-cells workgroup (fused dt)
  off = align_up(off + sizeof(double) * (size_t)std::max(std::max(1, 2 * nwg_of(p->pre_int)),
                                                         nwg_of(p->pre_all)));
  p->off_scal = off;
  off = align_up(off + sizeof(double) * 64);
  {
    int n1max = 1;
    for (const BlockDev& B : p->hb) n1max = std::max(n1max, B.n[1]);
    p->off_stat = off;
    off = align_up(off + sizeof(double) * (size_t)n1max * PLANE_CHUNKS * NSTAT);
  }
  p->ws_bytes = off;
  // device metadata
  std::vector<int> cdir(p->case_dir.begin(), p->case_dir.end());
  if (upload(p->ops, &p->d_ops) || upload(p->wall, &p->d_wall) || upload(p->tiles, &p->d_tiles) ||
      upload(p->pre_all, &p->d_pre_all) || upload(p->pre_box, &p->d_pre_box) ||
      upload(p->pre_int, &p->d_pre_int) || upload(p->pre_shell, &p->d_pre_shell) ||
      upload(p->case_wg, &p->d_case_wg) || upload(p->case_off, &p->d_case_off) ||
      upload(p->case_wg_all, &p->d_case_wg_all) || upload(p->case_off_all, &p->d_case_off_all) ||
      (!p->sponge_pre.empty() && upload(p->sponge_pre, &p->d_sponge_pre)) ||
      upload(cdir, &p->d_case_dir) || upload(p->case_target, &p->d_case_target)) {
    std::string m = g_err;
    wmx_plan_destroy(p);
    return fail(m);
  }
  if (hipMalloc((void**)&p->d_blks, sizeof(BlockDev) * nblocks) != hipSuccess) {
    wmx_plan_destroy(p);
    return fail("hipMalloc of block descriptors failed");
  }
  // the descriptors without workspace pointers, so that calls which read only the state
  // and the geometry (wmx_max_dt_noalloc) work before a workspace is bound
  if (hipMemcpy(p->d_blks, p->hb.data(), sizeof(BlockDev) * nblocks, hipMemcpyHostToDevice) != hipSuccess) {
    wmx_plan_destroy(p);
    return fail("upload of block descriptors failed");
  }
  *out = p;
  return 0;
}

int wmx_plan_destroy(wmx_plan* p) {
  if (!p) return 0;
  for (hipEvent_t e : p->ev_pool) (void)hipEventDestroy(e);
  for (auto& g : p->graphs) (void)hipGraphExecDestroy(g.exec);
  (void)hipFree(p->d_blks);
  (void)hipFree(p->d_ops);
  (void)hipFree(p->d_wall);
  (void)hipFree(p->d_tiles);
  (void)hipFree(p->d_pre_all);
  (void)hipFree(p->d_pre_box);
  (void)hipFree(p->d_pre_int);
  (void)hipFree(p->d_pre_shell);
  (void)hipFree(p->d_case_wg);
  (void)hipFree(p->d_case_off);
  (void)hipFree(p->d_case_wg_all);
  (void)hipFree(p->d_case_off_all);
  (void)hipFree(p->d_sponge_pre);
  (void)hipFree(p->d_case_dir);
  (void)hipFree(p->d_case_target);
  (void)hipFree(p->d_stage);
  (void)hipFree(p->d_pbuf);
  (void)hipFree(p->d_fill_off);
  (void)hipFree(p->d_fill_max);
  delete p;
  return 0;
}

int wmx_plan_workspace_bytes(const wmx_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail("null argument");
  *bytes = p->ws_bytes;
  return 0;
}

int wmx_plan_bind_workspace(wmx_plan* p, void* ws, size_t bytes) {
  if (!p || !ws) return fail("null argument");
  if (bytes < p->ws_bytes) return fail("workspace too small");
  if (reinterpret_cast<uintptr_t>(ws) % ALIGN) return fail("workspace must be 256-byte aligned");
  p->ws = static_cast<char*>(ws);
  for (int b = 0; b < p->nb; ++b) {
    BlockDev& B = p->hb[b];
    B.P = reinterpret_cast<double*>(p->ws + p->off_P[b]);
    B.G = reinterpret_cast<double*>(p->ws + p->off_G[b]);
    B.Pf = reinterpret_cast<float*>(p->ws + p->off_P[b]);
    B.Gf = reinterpret_cast<float*>(p->ws + p->off_G[b]);
    if (p->nd.prec == WMX_PREC_FP32_BULK) {
      float* g = reinterpret_cast<float*>(p->ws + p->off_geo32[b]);
      for (int d = 0; d < 3; ++d) {
        B.Sf[d] = g;
        g += 3 * fcount(B.N, d);
      }
      B.volf = g;
    } else {
      B.Sf[0] = B.Sf[1] = B.Sf[2] = nullptr;
      B.volf = nullptr;
    }
    B.Sb = reinterpret_cast<double*>(p->ws + p->off_sb[b]);
    B.shell = reinterpret_cast<double*>(p->ws + p->off_shell[b]);
    B.sponge_mean = reinterpret_cast<double*>(p->ws + p->off_sponge[b]);
    B.sponge_row = p->nd.n_sponge > 0
                       ? reinterpret_cast<int*>(B.sponge_mean + 5 * (size_t)B.n[0] * B.n[1])
                       : nullptr;
    for (int f = 0; f < 6; ++f)
      B.wall[f] = p->off_wall[b][f] >= 0 ? reinterpret_cast<double*>(p->ws + p->off_wall[b][f]) : nullptr;
  }
  p->d_part = reinterpret_cast<double*>(p->ws + p->off_part);
  p->d_delta = reinterpret_cast<double*>(p->ws + p->off_scal);
  p->d_stat = reinterpret_cast<double*>(p->ws + p->off_stat);
  WMX_CHECK(hipMemcpy(p->d_blks, p->hb.data(), sizeof(BlockDev) * p->nb, hipMemcpyHostToDevice));
  // the CFL dt's cell-mean face vectors, made once: the geometry may still be in flight on
  // the caller's stream, so the device is synchronised on both sides.  The planes that hold
  // +0.0 in every interior cell (axis-aligned directions) become BlockDev::sb_zero, whose
  // loads the dt skips; the kernel ORs the non-zero planes into one word per block, kept in
  // the partials buffer (at least two doubles per block, unused until the first step).
  WMX_CHECK(hipDeviceSynchronize());
  unsigned* d_nz = reinterpret_cast<unsigned*>(p->d_part);
  WMX_CHECK(hipMemset(d_nz, 0, sizeof(unsigned) * p->nb));
  for (int b = 0; b < p->nb; ++b) WMX_CHECK(launch_cell_sb(p->hb[b], d_nz + b, nullptr));
  WMX_CHECK(hipDeviceSynchronize());
  {
    std::vector<unsigned> nz(p->nb);
    WMX_CHECK(hipMemcpy(nz.data(), d_nz, sizeof(unsigned) * p->nb, hipMemcpyDeviceToHost));
    for (int b = 0; b < p->nb; ++b)
      p->hb[b].sb_zero = p->hb[b].Sb ? (int)(~nz[b] & ((1u << NSB_PLANES) - 1u)) : 0;
    WMX_CHECK(hipMemcpy(p->d_blks, p->hb.data(), sizeof(BlockDev) * p->nb, hipMemcpyHostToDevice));
  }
  if (p->nd.n_sponge > 0) {
    // rows of each block that need the sponge mean (static geometry; made once)
    WMX_CHECK(hipDeviceSynchronize());
    for (int b = 0; b < p->nb; ++b)
      WMX_CHECK(launch_sponge_rows(p->d_blks, b, (long long)p->hb[b].n[0] * p->hb[b].n[1], p->nd, nullptr));
    WMX_CHECK(hipDeviceSynchronize());
  }
  if (p->nd.prec == WMX_PREC_FP32_BULK) {
    // float geometry copies, made once: the geometry may still be in flight on the
    // caller's stream, so the device is synchronised on both sides of the copies
    WMX_CHECK(hipDeviceSynchronize());
    for (int b = 0; b < p->nb; ++b) {
      const BlockDev& B = p->hb[b];
      for (int d = 0; d < 3; ++d) WMX_CHECK(launch_to_f32(B.S[d], B.Sf[d], 3 * fcount(B.N, d), nullptr));
      WMX_CHECK(launch_to_f32(B.vol, B.volf, B.cells, nullptr));
    }
    WMX_CHECK(hipDeviceSynchronize());
  }
  return 0;
}

int wmx_plan_wall_faces(const wmx_plan* p, int32_t* count, int32_t* rec) {
  if (!p || !count) return fail("null argument");
  *count = (int32_t)p->wall.size();
  if (rec)
    for (size_t i = 0; i < p->wall.size(); ++i) {
      const WallOp& w = p->wall[i];
      int t1, t2;
      tangents(w.axis, &t1, &t2);
      rec[5 * i + 0] = w.blk;
      rec[5 * i + 1] = w.axis;
      rec[5 * i + 2] = w.side;
      rec[5 * i + 3] = p->hb[w.blk].n[t1];
      rec[5 * i + 4] = p->hb[w.blk].n[t2];
    }
  return 0;
}

int wmx_plan_fill_phases(const wmx_plan* p, int32_t* nphases) {
  if (!p || !nphases) return fail("null argument");
  *nphases = (int32_t)p->phase_max.size();
  return 0;
}

int wmx_plan_fill_passes(const wmx_plan* p, int32_t* npasses) {
  if (!p || !npasses) return fail("null argument");
  *npasses = (int32_t)p->passes;
  return 0;
}

int wmx_plan_set_case_gauge(wmx_plan* p, int32_t case_id, double p_ref, double T_ref) {
  if (!p) return fail("null plan");
  if (!(p_ref >= 0.0) || !(T_ref >= 0.0) || !(p_ref < HUGE_VAL) || !(T_ref < HUGE_VAL))
    return fail("gauge values must be finite and >= 0");
  if (case_id < -1 || case_id >= p->n_cases) return fail("case_id out of range");
  for (BlockDev& B : p->hb)
    if (case_id < 0 || B.case_id == case_id) {
      B.p_ref = p_ref;
      B.T_gauge = T_ref;
      B.e_ref = p->nd.gas.R * T_ref / (p->nd.gas.gamma - 1.0);
    }
  if (p->ws) {
    // the descriptors are read by kernels that may still be in flight on any stream
    WMX_CHECK(hipDeviceSynchronize());
    WMX_CHECK(hipMemcpy(p->d_blks, p->hb.data(), sizeof(BlockDev) * p->nb, hipMemcpyHostToDevice));
  }
  return 0;
}

int wmx_plan_set_gauge(wmx_plan* p, double p_ref, double T_ref) {
  return wmx_plan_set_case_gauge(p, -1, p_ref, T_ref);
}

int wmx_plan_set_wall_net(wmx_plan* p, const wmx_wall_net_desc* net) {
  if (p && net && p->persistent) return fail("disable persistent stages before setting a learned wall model");
  if (!p) return fail("null plan");
  if (net) {
    if (check_net(net)) return 1;
    p->nd.net = make_netdev(net);
  } else {
    std::memset(&p->nd.net, 0, sizeof(p->nd.net));
  }
  drop_graphs(p);
  return 0;
}

int wmx_plan_profile(wmx_plan* p, int32_t mask) {
  if (!p) return fail("null plan");
  p->prof_mask = mask;
  p->ev_used = 0;
  p->ev_rec.clear();
  return 0;
}

int wmx_plan_profile_read(wmx_plan* p, double* ms, int32_t* counts) {
  if (!p || !ms || !counts) return fail("null argument");
  for (int c = 0; c < K_NCLASS; ++c) { ms[c] = 0.0; counts[c] = 0; }
  for (auto& r : p->ev_rec) {
    hipEvent_t a = p->ev_pool[r.second], b = p->ev_pool[r.second + 1];
    WMX_CHECK(hipEventSynchronize(b));
    float t = 0.f;
    WMX_CHECK(hipEventElapsedTime(&t, a, b));
    ms[r.first] += t;
    counts[r.first] += 1;
  }
  return 0;
}

int wmx_fill_nodes(int32_t nblocks, const wmx_block_desc* blocks, double* const* X, int32_t ng,
                   int32_t node_passes, void* stream) {
  if (nblocks < 1 || nblocks > WMX_MAX_BLOCKS || !blocks || !X) return fail("bad arguments");
  if (ng != 3) return fail("ng must be 3");
  hipStream_t s = as_stream(stream);
  std::vector<BlockDev> hb(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    BlockDev& B = hb[b];
    std::memset(&B, 0, sizeof(B));
    for (int a = 0; a < 3; ++a) { B.n[a] = blocks[b].n[a]; B.N[a] = B.n[a] + 2 * ng; }
    B.ng = ng;
    B.cells = (long long)B.N[0] * B.N[1] * B.N[2];
    for (int f = 0; f < 6; ++f) B.faces[f] = blocks[b].faces[f];
  }
  BlockDev* d_blks = nullptr;
  GhostOp* d_ops = nullptr;
  std::vector<GhostOp> ops;
  for (int pass = 0; pass < node_passes; ++pass)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < nblocks; ++b)
        for (int sd = 0; sd < 2; ++sd) {
          int t1, t2;
          tangents(a, &t1, &t2);
          ops.push_back(GhostOp{b, a, sd, ng * (hb[b].N[t1] + 1) * (hb[b].N[t2] + 1)});
        }
  WMX_CHECK(hipMalloc((void**)&d_blks, sizeof(BlockDev) * nblocks));
  WMX_CHECK(hipMemcpy(d_blks, hb.data(), sizeof(BlockDev) * nblocks, hipMemcpyHostToDevice));
  if (upload(ops, &d_ops)) { (void)hipFree(d_blks); return 1; }
  Ptrs xp;
  std::memset(&xp, 0, sizeof(xp));
  for (int b = 0; b < nblocks; ++b) xp.p[b] = X[b];
  int rc = 0;
  for (size_t i = 0; i < ops.size() && !rc; ++i) {
    hipError_t e = launch_fill_nodes(d_blks, d_ops + i, 1, ops[i].count, xp, s);
    if (e != hipSuccess) rc = fail(std::string("node fill: ") + hipGetErrorString(e));
  }
  hipError_t e = hipStreamSynchronize(s);
  if (!rc && e != hipSuccess) rc = fail(std::string("node fill: ") + hipGetErrorString(e));
  (void)hipFree(d_blks);
  (void)hipFree(d_ops);
  return rc;
}

int wmx_compute_geometry(const double* X, const int32_t N[3], double* vol, double* xc, double* S0,
                         double* S1, double* S2, double* xf0, double* xf1, double* xf2,
                         void* stream) {
  if (!X || !N || !vol || !xc || !S0 || !S1 || !S2 || !xf0 || !xf1 || !xf2) return fail("null argument");
  int n3[3] = {N[0], N[1], N[2]};
  double* S[3] = {S0, S1, S2};
  double* xf[3] = {xf0, xf1, xf2};
  WMX_CHECK(launch_geometry(X, n3, vol, xc, S, xf, as_stream(stream)));
  return 0;
}

namespace {
int refuse_external(const wmx_plan* p) {
  if (p->has_external)
    return fail("the plan has faces the caller fills (WMX_BC_EXTERNAL): run the sweep with "
                "wmx_apply_bcs_phases and the RHS with wmx_rhs_filled");
  return 0;
}
}  // namespace

int wmx_apply_bcs(wmx_plan* p, double* const* q, double time, void* stream) {
  if (check_plan(p) || refuse_external(p)) return 1;
  return run_fill(p, make_ptrs(p, q), time, nullptr, 0.0, as_stream(stream));
}

int wmx_plan_sweep_phases(const wmx_plan* p, int32_t* nphases, int32_t* rec) {
  if (!p || !nphases) return fail("null argument");
  const int n = (int)p->phase_max.size();
  *nphases = n;
  if (rec)
    for (int i = 0; i < n; ++i) {
      const GhostOp& e = p->phase_ext[i];
      rec[4 * i] = e.blk >= 0 ? 1 : 0;
      rec[4 * i + 1] = e.blk;
      rec[4 * i + 2] = e.axis;
      rec[4 * i + 3] = e.side;
    }
  return 0;
}

int wmx_apply_bcs_phases(wmx_plan* p, double* const* q, int32_t first, int32_t last, double time,
                         void* stream) {
  if (check_plan(p)) return 1;
  if (!q) return fail("null argument");
  const int n = (int)p->phase_max.size();
  if (first < 0 || last > n || first > last) return fail("phase range out of bounds");
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q);
  for (int ph = first; ph < last; ++ph) {
    if (p->phase_ext[ph].blk >= 0) return fail("the phase range holds a caller-filled face");
    const int o = p->phase_off[ph], cnt = p->phase_off[ph + 1] - o;
    WMX_CHECK(launch_fill_cells(p->d_blks, p->d_ops + o, cnt, p->phase_max[ph], pq, p->nd.gas, s));
    for (int blk : p->phase_inflow[ph]) {
      const BlockDev& B = p->hb[blk];
      WMX_CHECK(launch_inflow(p->d_blks, blk, B.N[1], B.N[2], B.ng, p->inflows[B.faces[0].inflow_id], pq,
                              p->nd.gas, time, nullptr, 0.0, s));
    }
  }
  return 0;
}

int wmx_rhs_filled(wmx_plan* p, double* const* q, double* const* r, void* stream) {
  if (check_plan(p)) return 1;
  if (!q || !r) return fail("null argument");
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q), pr = make_ptrs(p, r);
  if (run_cells(p, pq, 0.0, nullptr, 0.0, s, nullptr, false, false)) return 1;
  WMX_CHECK(plan_flux(p, 0, p->d_blks, p->d_tiles, (int)p->tiles.size(), p->ci_len, p->nd, pq, pq, pr,
                        nullptr, 0.0, s));
  WMX_CHECK(launch_shell_finish(0, p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), pq, pr,
                                nullptr, 0.0, s));
  return 0;
}

int wmx_rhs(wmx_plan* p, double* const* q, double* const* r, double time, void* stream) {
  if (check_plan(p) || refuse_external(p)) return 1;
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q), pr = make_ptrs(p, r);
  if (run_cells(p, pq, time, nullptr, 0.0, s)) return 1;
  WMX_CHECK(plan_flux(p, 0, p->d_blks, p->d_tiles, (int)p->tiles.size(), p->ci_len, p->nd, pq, pq, pr,
                        nullptr, 0.0, s));
  WMX_CHECK(launch_shell_finish(0, p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), pq, pr,
                                nullptr, 0.0, s));
  return 0;
}

int wmx_mass_flux(wmx_plan* p, double* const* q, void* stream) {
  if (check_plan(p)) return 1;
  if (!p->any_constraint) return 0;
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q);
  WMX_CHECK(timed(p, K_MF, s, [&] {
    hipError_t e = launch_mf_partial(p->d_blks, p->d_pre_int, p->nb, nwg_of(p->pre_int), pq,
                                     p->d_case_dir, p->d_part, s);
    if (e == hipSuccess)
      e = launch_case_final(1, p->n_cases, p->case_wg_ident ? nullptr : p->d_case_wg, p->d_case_off, p->d_part,
                            p->d_case_target, p->d_delta, s);
    if (e == hipSuccess)
      e = launch_mf_apply(p->d_blks, p->d_pre_all, p->nb, nwg_of(p->pre_all), pq, p->d_case_dir,
                          p->d_delta, s);
    return e;
  }));
  return 0;
}

namespace {
// dt_out != nullptr (then dt_out == d_dt): the step's dt is computed here, from q0, by the
// stage-1 primitive launch.  The fill before it changes ghosts only, and stage 1's time is
// `time` itself (solver.py:125-129), so that fill takes no dt.
// Why a plan cannot run persistent stages ("" if it can).
std::string persistent_blocker(const wmx_plan* p) {
  if (p->nd.prec != WMX_PREC_FP32_BULK) return "persistent stages are built for the FP32-bulk form";
  for (auto& ph : p->phase_inflow)
    if (!ph.empty()) return "plans with a time-dependent inflow run separate launches";
  if (p->nd.n_sponge > 0) return "plans with sponge sources run separate launches";
  if (p->nd.net.w) return "plans with the learned wall model run separate launches";
  if (p->has_external) return "plans with caller-filled faces run separate launches";
  return "";
}

int upload_stage(wmx_plan* p) {
  if (!p->stage_dirty) return 0;
  StageStatic st;
  std::memset(&st, 0, sizeof(st));
  st.blks = p->d_blks;
  st.nd = p->nd;
  st.ops = p->d_ops;
  st.fill_off = p->d_fill_off;
  st.fill_max = p->d_fill_max;
  st.pre_all = p->d_pre_all;
  st.nblk = p->nb;
  st.nprim = nwg_of(p->pre_all);
  st.pre_box = p->d_pre_box;
  st.ngrad = nwg_of(p->pre_box);
  st.wall_ops = p->d_wall;
  st.nwall_ops = (int)p->wall.size();
  const int wx = st.nwall_ops > 0 && p->wall_max > 0 ? (p->wall_max + 255) / 256 : 0;
  st.grad_wall_wgx = p->fuse_grad ? 0 : wx;
  st.prim_wall_wgx = p->fuse_grad ? wx : 0;
  st.prim_head = p->fuse_grad ? (wx * st.nwall_ops + 7) / 8 * 8 : 0;
  if (p->fuse_grad) st.ngrad = 0;                 // the gradient list keeps only its dt head
  st.dt_list = p->case_wg_all_ident ? nullptr : p->d_case_wg_all;
  st.dt_off = p->d_case_off_all;
  st.tiles = p->d_tiles;
  st.ntiles = (int)p->tiles.size();
  st.ci_len = p->ci_len;
  st.pbuf = p->d_pbuf;
  WMX_CHECK(hipMemcpy(p->d_stage, &st, sizeof(st), hipMemcpyHostToDevice));
  p->stage_dirty = false;
  return 0;
}

// One step's stages as persistent launches (same work, same bits as the separate launches
// of advance_impl below): the ghost shell saved, three cooperative stage launches, the
// shell restored, the mass-flux constraint.
int advance_persistent(wmx_plan* p, double* const* q0, double* const* qout, double dt, const double* d_dt,
                       double* dt_out, void* stream) {
  if (p->capturing) return fail("persistent stages are not captured into graphs");
  if (upload_stage(p)) return 1;
  hipStream_t s = as_stream(stream);
  Ptrs P0 = make_ptrs(p, q0), PO = make_ptrs(p, qout);
  WMX_CHECK(launch_shell_save(p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), P0, s));
  const int sc = scheme_config(p->nd);
  const int nph = (int)p->phase_max.size();
  const int wx = !p->wall.empty() && p->wall_max > 0 ? (p->wall_max + 255) / 256 : 0;
  for (int mode = 1; mode <= 3; ++mode) {
    StageCall c;
    std::memset(&c, 0, sizeof(c));
    c.qin = mode == 1 ? P0 : PO;
    c.q0 = P0;
    c.out = PO;
    c.nfill = nph - (p->fuse_last ? 1 : 0);
    c.zfuse = p->fuse_last ? 1 : 0;
    c.ndt = mode == 1 && dt_out ? p->n_cases : 0;
    c.head = ((c.ndt + (p->fuse_grad ? 0 : wx * (int)p->wall.size())) + 7) / 8 * 8;
    c.part = c.ndt ? p->d_part : nullptr;
    c.dt_out = c.ndt ? dt_out : nullptr;
    c.d_dt = d_dt;
    c.dt_host = dt;
    WMX_CHECK(timed(p, K_FLUX, s, [&] {
      return launch_stage(mode, sc, p->fuse_grad, p->d_stage, c, p->fuse_grad ? &p->stage_grid_f : &p->stage_grid, s);
    }));
  }
  WMX_CHECK(launch_shell_finish(1, p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), P0, PO, d_dt, dt, s));
  return wmx_mass_flux(p, qout, stream);
}

int advance_impl(wmx_plan* p, double* const* q0, double* const* qout, double dt,
                 const double* d_dt, double* dt_out, double time, void* stream) {
  for (int b = 0; b < p->nb; ++b)
    if (q0[b] == qout[b]) return fail("qout must not alias q0");
  if (p->persistent) return advance_persistent(p, q0, qout, dt, d_dt, dt_out, stream);
  hipStream_t s = as_stream(stream);
  Ptrs P0 = make_ptrs(p, q0), PO = make_ptrs(p, qout);
  const int ntiles = (int)p->tiles.size();
  WMX_CHECK(timed(p, K_SHELL, s, [&] { return launch_shell_save(p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), P0, s); }));
  // stage times t, t + dt, t + dt/2 (solver.py:125-129)
  const double t1 = d_dt ? time : time + dt, t2 = d_dt ? time : time + 0.5 * dt;
  if (run_cells(p, P0, time, nullptr, 0.0, s, dt_out, true)) return 1;
  WMX_CHECK(timed(p, K_FLUX, s, [&] { return plan_flux(p, 1, p->d_blks, p->d_tiles, ntiles, p->ci_len, p->nd, P0, P0, PO, d_dt, dt, s); }));
  if (run_cells(p, PO, t1, d_dt, 1.0, s, nullptr, true)) return 1;
  WMX_CHECK(timed(p, K_FLUX, s, [&] { return plan_flux(p, 2, p->d_blks, p->d_tiles, ntiles, p->ci_len, p->nd, PO, P0, PO, d_dt, dt, s); }));
  if (run_cells(p, PO, t2, d_dt, 0.5, s, nullptr, true)) return 1;
  WMX_CHECK(timed(p, K_FLUX, s, [&] { return plan_flux(p, 3, p->d_blks, p->d_tiles, ntiles, p->ci_len, p->nd, PO, P0, PO, d_dt, dt, s); }));
  WMX_CHECK(timed(p, K_SHELL, s, [&] {
    return launch_shell_finish(1, p->d_blks, p->d_pre_shell, p->nb, nwg_of(p->pre_shell), P0, PO, d_dt, dt, s);
  }));
  return wmx_mass_flux(p, qout, stream);
}
}  // namespace

int wmx_advance(wmx_plan* p, double* const* q0, double* const* qout, double dt, const double* d_dt,
                double time, void* stream) {
  if (check_plan(p) || refuse_external(p)) return 1;
  return advance_impl(p, q0, qout, dt, d_dt, nullptr, time, stream);
}

int wmx_max_dt(wmx_plan* p, double* const* q, double* d_dt, void* stream) {
  if (check_plan(p)) return 1;
  if (!d_dt) return fail("null d_dt");
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q);
  WMX_CHECK(timed(p, K_DT, s, [&] {
    hipError_t e = launch_dt_partial(p->d_blks, p->d_pre_int, p->nb, nwg_of(p->pre_int), pq, p->nd,
                                     p->d_part, s);
    if (e == hipSuccess)
      e = launch_case_final(0, p->n_cases, p->case_wg_ident ? nullptr : p->d_case_wg, p->d_case_off,
                            p->d_part, nullptr, d_dt, s);
    return e;
  }));
  return 0;
}

// ---- allocation-free CFL dt -------------------------------------------------------
// The CFL partials of wmx_max_dt_noalloc live in this file-scope device array (one copy
// per GPU, held until the process ends).  Host code never passes the array's name to a
// kernel: its device address is looked up once per device (hipGetSymbolAddress, with the
// device current) and kept in g_scratch_ptr.  Every partial is written by the partial
// launch before the final launch reads it, so nothing needs clearing between calls.
static __device__ __align__(256) double scratch[WMX_NOALLOC_MAX_PARTIALS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_of_current_device(double** out) {
  int dev = -1;
  WMX_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device index out of range for the static scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    WMX_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail("hipGetSymbolAddress returned no address for the static scratch");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

int wmx_max_dt_noalloc(wmx_plan* p, double* const* q, double* d_dt, void* stream) {
  if (!p) return fail("null plan");
  if (!q || !d_dt) return fail("null argument");
  for (int b = 0; b < p->nb; ++b)
    if (!q[b]) return fail("null state pointer");
  const int npart = nwg_of(p->pre_int);
  if (npart > WMX_NOALLOC_MAX_PARTIALS)
    return fail("wmx_max_dt_noalloc: the plan needs " + std::to_string(npart) +
                " CFL partials, more than the static scratch holds (" +
                std::to_string(WMX_NOALLOC_MAX_PARTIALS) + ", i.e. about " +
                std::to_string((long long)WMX_NOALLOC_MAX_PARTIALS * 256 * INT_CELLS) +
                " interior cells); use wmx_max_dt");
  double* part = nullptr;
  if (scratch_of_current_device(&part)) return 1;
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q);
  // the same two launches as wmx_max_dt, with the partials in the static scratch
  WMX_CHECK(launch_dt_partial(p->d_blks, p->d_pre_int, p->nb, npart, pq, p->nd, part, s));
  WMX_CHECK(launch_case_final(0, p->n_cases, p->case_wg_ident ? nullptr : p->d_case_wg, p->d_case_off,
                              part, nullptr, d_dt, s));
  return 0;
}

int wmx_tile_model(int32_t nblocks, const int32_t* dims, int32_t precision, int32_t cu_count,
                   int32_t* ci_len, int32_t* ntiles) {
  if (nblocks < 1 || !dims || !ci_len || !ntiles) return fail("bad argument");
  long long T = 0;
  *ci_len = auto_ci_len(nblocks, dims, precision, cu_count, &T);
  *ntiles = (int32_t)T;
  return 0;
}

int wmx_plan_tiling(const wmx_plan* p, int32_t* ci_len, int32_t* order, int32_t* ntiles) {
  if (!p || !ci_len || !order || !ntiles) return fail("null argument");
  *ci_len = p->ci_len;
  *order = p->tile_order;
  *ntiles = (int32_t)p->tiles.size();
  return 0;
}

int wmx_plan_set_tiling(wmx_plan* p, int32_t ci_len, int32_t order) {
  if (!p) return fail("null plan");
  if (ci_len < 0 || ci_len > 1024) return fail("ci_len must be in [0 (auto), 1024]");
  if (order != 0 && order != 1) return fail("order must be 0 or 1");
  p->ci_req = ci_len;
  p->tile_order = order;
  build_tiles(p);
  (void)hipFree(p->d_tiles);
  p->d_tiles = nullptr;
  if (upload(p->tiles, &p->d_tiles)) return 1;
  drop_graphs(p);
  p->stage_dirty = true;
  return 0;
}

int wmx_plan_set_fused_gradients(wmx_plan* p, int32_t on) {
  if (check_plan(p)) return 1;
  p->fuse_grad = on != 0;
  p->stage_dirty = true;
  drop_graphs(p);
  return 0;
}

int wmx_plan_set_persistent(wmx_plan* p, int32_t on) {
  if (check_plan(p)) return 1;
  if (!on) {
    p->persistent = false;
    return 0;
  }
  const std::string why = persistent_blocker(p);
  if (!why.empty()) return fail(why);
  if (!p->d_stage) {
    std::vector<int> moff(p->phase_off.begin(), p->phase_off.end());
    std::vector<int> mmax(p->phase_max.begin(), p->phase_max.end());
    if (mmax.empty()) mmax.push_back(0);
    WMX_CHECK(hipMalloc((void**)&p->d_stage, sizeof(StageStatic)));
    WMX_CHECK(hipMalloc((void**)&p->d_pbuf, 3 * sizeof(Ptrs)));
    if (upload(moff, &p->d_fill_off) || upload(mmax, &p->d_fill_max)) return 1;
  }
  drop_graphs(p);
  p->stage_dirty = true;
  p->persistent = true;
  return 0;
}

int wmx_step(wmx_plan* p, double* const* q0, double* const* qout, double* d_dt, double time,
             int32_t use_graph, void* stream) {
  if (check_plan(p) || refuse_external(p)) return 1;
  if (!d_dt) return fail("null d_dt");
  hipStream_t s = as_stream(stream);
  bool has_inflow = false;
  for (auto& ph : p->phase_inflow) has_inflow |= !ph.empty();
  for (int b = 0; b < p->nb; ++b)
    if (q0[b] == qout[b]) return fail("qout must not alias q0");
  // the step's dt comes from its stage-1 primitive launch; time-dependent inflow plans
  // run eagerly (a captured graph would hold one time)
  if (!use_graph || has_inflow || s == nullptr || p->persistent)
    return advance_impl(p, q0, qout, 0.0, d_dt, d_dt, time, stream);
  for (auto& g : p->graphs) {
    bool same = g.d_dt == d_dt && g.time == time;
    for (int b = 0; same && b < p->nb; ++b) same = g.q0[b] == q0[b] && g.qout[b] == qout[b];
    if (same) {
      WMX_CHECK(hipGraphLaunch(g.exec, s));
      return 0;
    }
  }
  hipGraph_t graph;
  WMX_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  p->capturing = true;
  int rc = advance_impl(p, q0, qout, 0.0, d_dt, d_dt, time, stream);
  p->capturing = false;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) return 1;
  if (e != hipSuccess) return fail(std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t exec;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(std::string("graph instantiate: ") + hipGetErrorString(e));
  if (p->graphs.size() >= 4) {
    (void)hipGraphExecDestroy(p->graphs.front().exec);
    p->graphs.erase(p->graphs.begin());
  }
  wmx_plan::GraphEntry ge;
  ge.q0.assign(q0, q0 + p->nb);
  ge.qout.assign(qout, qout + p->nb);
  ge.d_dt = d_dt;
  ge.time = time;
  ge.exec = exec;
  p->graphs.push_back(ge);
  WMX_CHECK(hipGraphLaunch(exec, s));
  return 0;
}

int wmx_wall_stats(wmx_plan* p, double* const* q, double* d_tau, double* d_qw, void* stream) {
  if (check_plan(p)) return 1;
  hipStream_t s = as_stream(stream);
  Ptrs pq = make_ptrs(p, q);
  WMX_CHECK(launch_wall(p->d_blks, p->d_wall, (int)p->wall.size(), p->wall_max, pq, p->nd, s));
  size_t off = 0;
  for (const WallOp& w : p->wall) {
    const double* W = p->hb[w.blk].wall[2 * w.axis + w.side];
    WMX_CHECK(hipMemcpyAsync(d_tau + off, W + 5 * (size_t)w.count, sizeof(double) * w.count,
                             hipMemcpyDeviceToDevice, s));
    WMX_CHECK(hipMemcpyAsync(d_qw + off, W + 6 * (size_t)w.count, sizeof(double) * w.count,
                             hipMemcpyDeviceToDevice, s));
    off += w.count;
  }
  return 0;
}

int wmx_plane_stats(wmx_plan* p, double* const* q, int32_t block, double* d_acc, void* stream) {
  if (check_plan(p)) return 1;
  if (!q || !d_acc) return fail("null argument");
  if (block < 0 || block >= p->nb) return fail("plane stats: block out of range");
  hipStream_t s = as_stream(stream);
  WMX_CHECK(launch_plane_stats(p->d_blks, p->hb[block], block, make_ptrs(p, q), p->nd.gas, p->d_stat,
                               d_acc, s));
  return 0;
}

int wmx_fill_face(const wmx_block_desc* blk, int32_t axis, int32_t side, const wmx_numerics_desc* nm,
                  double* q, double time, void* stream) {
  (void)time;
  if (!blk || !nm || !q) return fail("null argument");
  if (axis < 0 || axis > 2 || side < 0 || side > 1) return fail("bad face");
  const wmx_face_desc& f = blk->faces[2 * axis + side];
  if (f.kind == WMX_FACE_INTERFACE || f.kind == WMX_BC_INFLOW || f.kind < 0 || f.kind > WMX_BC_INFLOW)
    return fail("wmx_fill_face handles boundary-condition faces only");
  const int ng = 3;
  BlockDev B;
  std::memset(&B, 0, sizeof(B));
  for (int a = 0; a < 3; ++a) { B.n[a] = blk->n[a]; B.N[a] = blk->n[a] + 2 * ng; }
  B.ng = ng;
  B.cells = (long long)B.N[0] * B.N[1] * B.N[2];
  B.vol = blk->vol;
  B.xc = blk->xc;
  for (int a = 0; a < 3; ++a) { B.S[a] = blk->S[a]; B.xf[a] = blk->xf[a]; }
  for (int k = 0; k < 6; ++k) B.faces[k] = blk->faces[k];
  if (!B.S[axis]) return fail("face vectors required");
  int t1, t2;
  tangents(axis, &t1, &t2);
  GhostOp op{0, axis, side, ng * B.N[t1] * B.N[t2]};
  hipStream_t s = as_stream(stream);
  BlockDev* d_b = nullptr;
  GhostOp* d_op = nullptr;
  WMX_CHECK(hipMalloc((void**)&d_b, sizeof(BlockDev)));
  WMX_CHECK(hipMalloc((void**)&d_op, sizeof(GhostOp)));
  WMX_CHECK(hipMemcpy(d_b, &B, sizeof(BlockDev), hipMemcpyHostToDevice));
  WMX_CHECK(hipMemcpy(d_op, &op, sizeof(GhostOp), hipMemcpyHostToDevice));
  Ptrs pq;
  std::memset(&pq, 0, sizeof(pq));
  pq.p[0] = q;
  NumDev nd = make_numdev(nm);
  hipError_t e = launch_fill_cells(d_b, d_op, 1, op.count, pq, nd.gas, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_b);
  (void)hipFree(d_op);
  if (e != hipSuccess) return fail(std::string("fill_face: ") + hipGetErrorString(e));
  return 0;
}

int wmx_wall_model_eval(const wmx_numerics_desc* nm, int32_t mode, int64_t n, const double* a,
                        const double* b, const double* c, const double* d, double T_wall,
                        double* o0, double* o1, void* stream) {
  if (!nm || !a || !b || !c || !d || !o0 || (mode == 1 && !o1)) return fail("null argument");
  if (mode != 0 && mode != 1) return fail("mode must be 0 or 1");
  NumDev nd = make_numdev(nm);
  WMX_CHECK(launch_wm_eval(nd, mode, n, a, b, c, d, T_wall, o0, o1, as_stream(stream)));
  return 0;
}

int wmx_mlp_sizes(int32_t* nparam, int32_t* nbf16, int32_t* nfp32) {
  if (!nparam || !nbf16 || !nfp32) return fail("null argument");
  *nparam = mlp_nparam();
  *nbf16 = mlp_nbf16();
  *nfp32 = mlp_nf32();
  return 0;
}

int wmx_wm_dataset(const wmx_numerics_desc* nm, const wmx_wall_net_desc* net, int64_t n, uint64_t seed,
                   float* X, float* Y, double* raw, void* stream) {
  if (!nm || !net || !X || !Y) return fail("null argument");
  if (n < 0) return fail("n must be >= 0");
  for (int k = 0; k < 3; ++k)
    if (!(net->hi[k] >= net->lo[k]) || !(net->x_scale[k] > 0.f)) return fail("wall net: bad input range");
  for (int k = 0; k < 2; ++k)
    if (!(net->y_scale[k] > 0.f)) return fail("wall net: y_scale must be > 0");
  NumDev nd = make_numdev(nm);
  WallNetDev d;
  std::memset(&d, 0, sizeof(d));
  for (int k = 0; k < 3; ++k) {
    d.xc[k] = net->x_center[k];
    d.xs[k] = net->x_scale[k];
    d.lo[k] = net->lo[k];
    d.hi[k] = net->hi[k];
  }
  for (int k = 0; k < 2; ++k) {
    d.yc[k] = net->y_center[k];
    d.ys[k] = net->y_scale[k];
  }
  nd.net = d;
  WMX_CHECK(launch_wm_dataset(nd, n, seed, X, Y, raw, as_stream(stream)));
  return 0;
}

int wmx_mlp_train_step(float* params, const float* X, const float* Y, int64_t n_samples, int64_t offset,
                       int32_t batch, float* partials, float* loss_partials, float* grad, float* adam_m,
                       float* adam_v, int32_t step, float lr, int32_t update, void* stream) {
  if (!params || !X || !Y || !partials || !loss_partials) return fail("null argument");
  if (batch < 1) return fail("batch must be >= 1");
  if (offset < 0 || offset + batch > n_samples) return fail("batch [offset, offset + batch) exceeds the samples");
  if (update && (!adam_m || !adam_v || step < 1)) return fail("update needs adam_m, adam_v and step >= 1");
  WMX_CHECK(launch_mlp_train(params, X, Y, offset, batch, partials, loss_partials, grad, adam_m, adam_v,
                             step, lr, update, as_stream(stream)));
  return 0;
}

int wmx_mlp_pack(const float* params, uint16_t* w_bf16, float* w_fp32, void* stream) {
  if (!params || !w_bf16 || !w_fp32) return fail("null argument");
  WMX_CHECK(launch_mlp_pack(params, w_bf16, w_fp32, as_stream(stream)));
  return 0;
}

int wmx_mlp_forward(const float* params, const float* X, int64_t n, float* Y, void* stream) {
  if (!params || !X || !Y) return fail("null argument");
  if (n < 0) return fail("n must be >= 0");
  WMX_CHECK(launch_mlp_forward32(params, X, n, Y, as_stream(stream)));
  return 0;
}

int wmx_wall_net_eval(const wmx_numerics_desc* nm, const wmx_wall_net_desc* net, int32_t mode, int64_t n,
                      const double* h, const double* hdel, const double* U, const double* T_h,
                      const double* p_h, double T_wall, double* tau_w, double* q_w, double* uplus,
                      void* stream) {
  if (!nm || !net || !h || !U || !T_h || !p_h || !tau_w || !q_w) return fail("null argument");
  if (mode != 0 && mode != 1) return fail("mode must be 0 or 1");
  if (n < 0) return fail("n must be >= 0");
  if (check_net(net)) return 1;
  NumDev nd = make_numdev(nm);
  nd.net = make_netdev(net);
  WMX_CHECK(launch_wm_net_eval(nd, mode, n, h, hdel, U, T_h, p_h, T_wall, tau_w, q_w, uplus,
                               as_stream(stream)));
  return 0;
}

int wmx_scheme_flux(const wmx_numerics_desc* nm, const int32_t n[3], int32_t d, const double* w,
                    const double* S_d, const double* theta, double* F, void* stream) {
  if (!nm || !n || !w || !S_d || !F) return fail("null argument");
  if (d < 0 || d > 2) return fail("direction must be 0, 1 or 2");
  if (n[0] < 3 || n[1] < 3 || n[2] < 3) return fail("each direction needs at least ng = 3 cells");
  if (nm->scheme == WMX_SCHEME_HYBRID && !theta) return fail("the Hybrid scheme needs theta");
  if (nm->scheme < WMX_SCHEME_KEP || nm->scheme > WMX_SCHEME_HYBRID) return fail("unknown scheme");
  if ((nm->scheme != WMX_SCHEME_MUSCLHLLC) && nm->order != 2 && nm->order != 4) return fail("order must be 2 or 4");
  NumDev nd = make_numdev(nm);
  WMX_CHECK(launch_scheme_flux(nd, n, d, w, S_d, theta, F, as_stream(stream)));
  return 0;
}

int wmx_sgs_eval(const wmx_numerics_desc* nm, int64_t n, const double* g, const double* delta,
                 const double* rho, double* mu_t, void* stream) {
  if (!nm || !g || !delta || !rho || !mu_t) return fail("null argument");
  if (n < 0) return fail("n must be >= 0");
  if (nm->sgs < WMX_SGS_NONE || nm->sgs > WMX_SGS_VREMAN) return fail("unknown SGS model");
  NumDev nd = make_numdev(nm);
  WMX_CHECK(launch_sgs_eval(nd, n, g, delta, rho, mu_t, as_stream(stream)));
  return 0;
}

int wmx_wall_flux(const wmx_block_desc* blk, int32_t axis, int32_t side, const wmx_numerics_desc* nm,
                  const double* w, const double* T, double* F, double* tau_w, double* q_w, void* stream) {
  if (!blk || !nm || !w || !T || !F || !tau_w || !q_w) return fail("null argument");
  if (axis < 0 || axis > 2 || side < 0 || side > 1) return fail("bad face");
  const wmx_face_desc& f = blk->faces[2 * axis + side];
  if (f.kind != WMX_BC_WALLMODEL) return fail("wmx_wall_flux needs a WallModel face");
  if (f.n_match < 1 || f.n_match > blk->n[axis]) return fail("n_match out of range");
  if (!blk->S[axis] || !blk->xf[axis] || !blk->xc) return fail("face vectors, face centres and centroids required");
  const int ng = 3;
  BlockDev B;
  std::memset(&B, 0, sizeof(B));
  for (int a = 0; a < 3; ++a) { B.n[a] = blk->n[a]; B.N[a] = blk->n[a] + 2 * ng; }
  B.ng = ng;
  B.cells = (long long)B.N[0] * B.N[1] * B.N[2];
  B.vol = blk->vol;
  B.xc = blk->xc;
  for (int a = 0; a < 3; ++a) { B.S[a] = blk->S[a]; B.xf[a] = blk->xf[a]; }
  for (int k = 0; k < 6; ++k) B.faces[k] = blk->faces[k];
  NumDev nd = make_numdev(nm);
  WMX_CHECK(launch_wall_flux(nd, B, axis, side, w, T, F, tau_w, q_w, as_stream(stream)));
  return 0;
}

int wmx_mfma_bf16_probe(const float* A, const float* B, float* D, void* stream) {
  if (!A || !B || !D) return fail("null argument");
  WMX_CHECK(launch_mfma_probe(A, B, D, as_stream(stream)));
  return 0;
}

}  // extern "C"

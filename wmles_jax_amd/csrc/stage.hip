// One SSP-RK3 stage of an FP32-bulk plan in a single cooperative launch: a persistent grid
// (the co-resident maximum, a multiple of 8) walks the stage's work lists in order -- the
// ghost sweep's phases, the primitive workgroups (stage 1: with the CFL partials), the
// gradient workgroups with their head work (stage 1: the per-case dt mins; the Newton
// wall faces) and the flux tiles -- with a grid-wide barrier after each phase (each is read
// by the next: ghosts by the primitives, records by the gradients, records and wall fluxes
// by the flux tiles; cooperative_groups' grid barrier writes back and invalidates the
// XCDs' L2 caches, so the phases see each other's stores).  Work item w of a phase is
// taken by workgroup w mod grid, so with a grid that is a multiple of 8 it runs on XCD
// w mod 8 as in the separate launches, whose XCD-span orderings (xcd_span, the tile order)
// carry over.  The bodies are the separate kernels' own (stage_off.hip / stage_fast.hip,
// device-linked with their units' FP contraction), so the results are the same bits.
#include <hip/hip_cooperative_groups.h>

#include "common.h"

namespace cg = cooperative_groups;

namespace wmx {

__device__ void st_fill(const StageStatic& s, int opi, int idx);
template <bool DT>
__device__ void st_prim(const StageStatic& s, double* __restrict__ part, int zfuse, int wgx);
__device__ void st_grad(const StageStatic& s, int head, int ndt, const double* __restrict__ part,
                        double* __restrict__ dt_out, int wgx);
template <int MODE, int SC>
__device__ void st_flux(const StageStatic& s, const double* __restrict__ d_dt, double dt_host, int tile);
template <int MODE, int SC>
__device__ void st_flux3(const StageStatic& s, const double* __restrict__ d_dt, double dt_host, int tile);

// FUSED: the fused kernel form (flux3 tiles with the gradient records in the march, wall
// faces in the primitive head, the gradient list only its dt head)
template <int MODE, int SC, bool FUSED>
__global__ __launch_bounds__(256, 2) void k_stage(const StageStatic* __restrict__ st, StageCall c) {
  cg::grid_group grid = cg::this_grid();
  const StageStatic& s = *st;
  const int G = (int)gridDim.x;
  // the launch's state pointers, where the bodies index them
  if (blockIdx.x == 0 && threadIdx.x < WMX_MAX_BLOCKS) {
    const int t = (int)threadIdx.x;
    s.pbuf[0].p[t] = c.qin.p[t];
    s.pbuf[1].p[t] = c.q0.p[t];
    s.pbuf[2].p[t] = c.out.p[t];
  }
  grid.sync();
  for (int ph = 0; ph < c.nfill; ++ph) {
    const int o = s.fill_off[ph], nops = s.fill_off[ph + 1] - o;
    const int per = (s.fill_max[ph] + 255) >> 8;             // workgroups per op
    const int nwork = nops * per;
    for (int w = (int)blockIdx.x; w < nwork; w += G) st_fill(s, o + w / per, (w % per) * 256 + (int)threadIdx.x);
    grid.sync();
  }
  const int npr = s.prim_head + s.nprim;
  for (int w = (int)blockIdx.x; w < npr; w += G) {
    if (c.part) st_prim<true>(s, c.part, c.zfuse, w);          // (uniform: a kernel argument)
    else st_prim<false>(s, c.part, c.zfuse, w);
    __syncthreads();                                            // LDS reuse by the next item
  }
  grid.sync();
  const int ng = c.head + s.ngrad;
  for (int w = (int)blockIdx.x; w < ng; w += G) {
    st_grad(s, c.head, c.ndt, c.part, c.dt_out, w);
    __syncthreads();
  }
  grid.sync();
  for (int w = (int)blockIdx.x; w < s.ntiles; w += G) {
    if (FUSED) st_flux3<MODE, SC>(s, c.d_dt, c.dt_host, w);
    else st_flux<MODE, SC>(s, c.d_dt, c.dt_host, w);
    __syncthreads();
  }
}

namespace {
template <int MODE, int SC, bool FUSED>
hipError_t coop(const StageStatic* d_st, const StageCall& c, int grid, int* used, hipStream_t s) {
  const void* fn = (const void*)k_stage<MODE, SC, FUSED>;
  if (grid <= 0) {
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0);
    if (e != hipSuccess) return e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    grid = (per_cu * cus) / 8 * 8;
    if (grid < 8) return hipErrorInvalidConfiguration;
  }
  if (used) *used = grid;
  const StageStatic* a0 = d_st;
  StageCall a1 = c;
  void* args[] = {(void*)&a0, (void*)&a1};
  return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), args, 0, s);
}
}  // namespace

hipError_t launch_stage(int mode, int sc, bool fused, const StageStatic* d_st, const StageCall& c,
                        int* grid_used, hipStream_t s) {
  const int grid = grid_used && *grid_used > 0 ? *grid_used : 0;
#define WMX_STAGE_F(M, C) \
  return fused ? coop<M, C, true>(d_st, c, grid, grid_used, s) : coop<M, C, false>(d_st, c, grid, grid_used, s);
#define WMX_STAGE_M(C)                 \
  switch (mode) {                      \
    case 1: WMX_STAGE_F(1, C)          \
    case 2: WMX_STAGE_F(2, C)          \
    default: WMX_STAGE_F(3, C)         \
  }
  switch (sc) {
    case SC_HYB_KEP4_ALBADA: WMX_STAGE_M(SC_HYB_KEP4_ALBADA)
    case SC_KEP2: WMX_STAGE_M(SC_KEP2)
    case SC_KEP4: WMX_STAGE_M(SC_KEP4)
    default: WMX_STAGE_M(SC_ANY)
  }
#undef WMX_STAGE_M
#undef WMX_STAGE_F
}

}  // namespace wmx

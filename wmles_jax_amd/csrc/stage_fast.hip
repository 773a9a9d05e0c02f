// Bodies of the persistent stage kernel (stage.hip) from the contracted units: the
// FP32-bulk gradients (grad32.h) and flux tiles (flux2.h), built with the same FP
// contraction as rhs.hip / flux.hip and as relocatable device code, so the stage kernel's
// records and stage results are the same bits as the separate launches'.
#include "flux3.h"
#include "grad32.h"

namespace wmx {

__device__ void st_grad(const StageStatic& s, int head, int ndt, const double* __restrict__ part,
                        double* __restrict__ dt_out, int wgx) {
  const int nwg = head + s.ngrad;
#define WMX_ST_GRAD(S)                                                                                     \
  grad32_block<S>(s.blks, s.pre_box, s.nblk, s.nd, head, s.wall_ops, s.grad_wall_wgx, s.nwall_ops, s.pbuf[0], \
                  ndt, s.dt_list, s.dt_off, part, dt_out, wgx, nwg)
  switch (s.nd.sgs) {
    case WMX_SGS_NONE: WMX_ST_GRAD(WMX_SGS_NONE); break;
    case WMX_SGS_SMAGORINSKY: WMX_ST_GRAD(WMX_SGS_SMAGORINSKY); break;
    case WMX_SGS_WALE: WMX_ST_GRAD(WMX_SGS_WALE); break;
    default: WMX_ST_GRAD(WMX_SGS_VREMAN);
  }
#undef WMX_ST_GRAD
}

template <int MODE, int SC>
__device__ void st_flux(const StageStatic& s, const double* __restrict__ d_dt, double dt_host, int tile) {
  flux2_tile<float, MODE, SC>(s.blks, s.tiles, s.ci_len, s.nd, s.pbuf[0], s.pbuf[1], s.pbuf[2], d_dt, dt_host,
                              tile);
}

template <int MODE, int SC>
__device__ void st_flux3(const StageStatic& s, const double* __restrict__ d_dt, double dt_host, int tile) {
  flux3_tile<MODE, SC>(s.blks, s.tiles, s.ci_len, s.nd, s.pbuf[0], s.pbuf[1], s.pbuf[2], d_dt, dt_host, tile);
}

#define WMX_ST_FLUX(M, C)                                                          \
  template __device__ void st_flux<M, C>(const StageStatic&, const double*, double, int); \
  template __device__ void st_flux3<M, C>(const StageStatic&, const double*, double, int);
#define WMX_ST_FLUX_M(C) WMX_ST_FLUX(1, C) WMX_ST_FLUX(2, C) WMX_ST_FLUX(3, C)
WMX_ST_FLUX_M(SC_ANY)
WMX_ST_FLUX_M(SC_HYB_KEP4_ALBADA)
WMX_ST_FLUX_M(SC_KEP2)
WMX_ST_FLUX_M(SC_KEP4)
#undef WMX_ST_FLUX_M
#undef WMX_ST_FLUX

}  // namespace wmx

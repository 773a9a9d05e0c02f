// Launches of the fused FP32-bulk step kernel whose tiles also evaluate the gradient
// records (flux3.h), one instance per stage mode and scheme configuration.
#include "flux3.h"

namespace wmx {

template <int MODE, int SC>
__global__ __launch_bounds__(256, 2) void k_flux3(const BlockDev* __restrict__ blks,
                                               const int4* __restrict__ tiles, int ci_len,
                                               NumDev nd, Ptrs qin, Ptrs q0, Ptrs out,
                                               const double* __restrict__ d_dt, double dt_host) {
  flux3_tile<MODE, SC>(blks, tiles, ci_len, nd, qin, q0, out, d_dt, dt_host, (int)blockIdx.x);
}

hipError_t launch_flux3(int mode, const BlockDev* d_blks, const int4* d_tiles, int ntiles, int ci_len,
                        const NumDev& nd, Ptrs qin, Ptrs q0, Ptrs out, const double* d_dt, double dt,
                        hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  dim3 grid(ntiles), block(NT);
#define WMX_F3(M, C) hipLaunchKernelGGL((k_flux3<M, C>), grid, block, 0, s, d_blks, d_tiles, ci_len, nd, qin, q0, out, d_dt, dt)
#define WMX_F3M(C)              \
  switch (mode) {               \
    case 0: WMX_F3(0, C); break; \
    case 1: WMX_F3(1, C); break; \
    case 2: WMX_F3(2, C); break; \
    default: WMX_F3(3, C); break; \
  }
  switch (scheme_config(nd)) {
    case SC_HYB_KEP4_ALBADA: WMX_F3M(SC_HYB_KEP4_ALBADA); break;
    case SC_KEP2: WMX_F3M(SC_KEP2); break;
    case SC_KEP4: WMX_F3M(SC_KEP4); break;
    default: WMX_F3M(SC_ANY); break;
  }
#undef WMX_F3M
#undef WMX_F3
  return hipGetLastError();
}

}  // namespace wmx

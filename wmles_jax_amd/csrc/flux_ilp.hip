// The FP64 flux kernels and the Hybrid (KEP4 + MUSCL-HLLC) FP32-bulk flux kernels of
// flux.hip, compiled in a unit of their own so that the build can give them LLVM's max-ILP
// machine scheduler; flux.hip declares them extern and launches them.
#define WMX_FLUX_ILP_UNIT 1
#include "flux.hip"

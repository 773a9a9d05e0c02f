// Bodies of the persistent stage kernel (stage.hip) that must round as the ghost-fill and
// primitive units do: built, like them, without FP contraction and as relocatable device
// code, so the stage kernel's ghost values, primitive records and CFL partials are the same
// bits as the separate launches'.
#include "prim32.h"

namespace wmx {

__device__ void st_fill(const StageStatic& s, int opi, int idx) {
  fill_cell(s.blks, s.ops[opi], s.pbuf[0], s.nd.gas, idx);
}

template <bool DT>
__device__ void st_prim(const StageStatic& s, double* __restrict__ part, int zfuse, int wgx) {
  SpongeHead sh;                               // (no sponge rows in a persistent plan)
  sh.wall = s.wall_ops;
  sh.wall_wgx = s.prim_wall_wgx;
  sh.nwall = s.prim_wall_wgx > 0 ? s.nwall_ops : 0;
  sh.head = s.prim_head;
  prim32_block<DT>(s.blks, s.pre_all, s.nblk, s.pbuf[0], s.nd, part, zfuse, sh, wgx, s.prim_head + s.nprim);
}
template __device__ void st_prim<true>(const StageStatic&, double*, int, int);
template __device__ void st_prim<false>(const StageStatic&, double*, int, int);

}  // namespace wmx

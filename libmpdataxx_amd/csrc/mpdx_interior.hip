// The FP32 interior kernels with the default guards (the production FP32
// path), compiled as a translation unit of their own so that they are built
// with the max-ILP instruction scheduler: their occupancy is fixed by LDS
// (one 1024-thread block per CU, or two 512-thread ones), so registers are
// better spent on instruction-level parallelism than on the occupancy the
// default scheduler aims for; the 24-row tiles of the later corrective
// passes are compiled for two 768-thread blocks per CU (6 waves per SIMD,
// interior_waves_per_eu).  Both the single launches (k_fz_interior) and
// the batched ones of ensembles (k_task_list over InteriorTask and
// InteriorMixTask) are here.
// The FP64 and other-guard variants keep the default scheduler (with
// max-ILP they need up to 250 VGPRs or spill).
#define MPDX_INTERIOR_TU 1
#include "mpdx_tasks.hpp"
#include "mpdx_interior_list.hpp"

namespace mpdx {
#define MPDX_INST(R, G, F, I, S, Z, T, W)                                                     \
  template __global__ void k_fz_interior<R, G, F, I, S, Z, T, W>(InteriorArgs<R>);            \
  template __global__ void k_task_list<InteriorTask<R, G, F, I, S, Z, T, W>>(                 \
      const InteriorTask<R, G, F, I, S, Z, T, W> *, const unsigned *, int);
MPDX_INTERIOR_F32_FAST(MPDX_INST)
#undef MPDX_INST
#define MPDX_INST_MIX(R, G, T, W)                                                             \
  template __global__ void k_task_list<InteriorMixTask<R, G, T, W>>(const InteriorMixTask<R, G, T, W> *, \
                                                                    const unsigned *, int);
MPDX_INTERIOR_MIX_F32_FAST(MPDX_INST_MIX)
#undef MPDX_INST_MIX
}  // namespace mpdx

// The FP32 interior kernels with the default guards (the production FP32
// path), compiled as a translation unit of their own so that they are built
// with the max-ILP instruction scheduler: their occupancy is fixed by LDS
// (one 1024-thread block per CU), so registers are better spent on
// instruction-level parallelism than on the occupancy the default scheduler
// aims for.  The FP64 and other-guard variants keep the default scheduler
// (with max-ILP they need up to 250 VGPRs or spill).
#include "mpdx_fused.hpp"
#include "mpdx_interior_list.hpp"

namespace mpdx {
#define MPDX_INST(R, G, F, I, S, Z) template __global__ void k_fz_interior<R, G, F, I, S, Z>(InteriorArgs<R>);
MPDX_INTERIOR_F32_FAST(MPDX_INST)
#undef MPDX_INST
}  // namespace mpdx

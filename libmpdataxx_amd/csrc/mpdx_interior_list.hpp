// The FP32 interior kernels with the default (fast FP64-epsilon) guards:
// k_fz_interior<float, GuardFast, FCT, IGA, STOREV, FUSE0> for every option
// combination the library launches.  X(R, G, FCT, IGA, STOREV, FUSE0) is
// applied to each.
#pragma once

#define MPDX_IV_S(X, F, I, Z)                                         \
  X(float, ::mpdx::GuardFast, F, I, true, Z)                          \
  X(float, ::mpdx::GuardFast, F, I, false, Z)

#define MPDX_IV_Z(X, F, I) MPDX_IV_S(X, F, I, true) MPDX_IV_S(X, F, I, false)

#define MPDX_INTERIOR_F32_FAST(X)                                     \
  MPDX_IV_Z(X, true, false)                                           \
  MPDX_IV_Z(X, true, true)                                            \
  MPDX_IV_Z(X, false, false)                                          \
  MPDX_IV_Z(X, false, true)

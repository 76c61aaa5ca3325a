// The FP32 interior kernels with the default (fast FP64-epsilon) guards:
// k_fz_interior<float, GuardFast, FCT, IGA, STOREV, FUSE0, TY, WXY> for
// every option combination the library launches, both tile heights (32,
// 16) and both x / y bound treatments (WXY: wrapped to the periodic images),
// and 24-row tiles for the corrective passes after the first (FUSE0 false)
// without iga (an iga step has a single corrective pass).
// X(R, G, FCT, IGA, STOREV, FUSE0, TY, WXY) is applied to each.
#pragma once

#define MPDX_IV_24_true_true(X, F, S)
#define MPDX_IV_24_true_false(X, F, S)
#define MPDX_IV_24_false_true(X, F, S)
#define MPDX_IV_24_false_false(X, F, S)                               \
  X(float, ::mpdx::GuardFast, F, false, S, false, 24, false)          \
  X(float, ::mpdx::GuardFast, F, false, S, false, 24, true)

#define MPDX_IV_T(X, F, I, S, Z)                                      \
  X(float, ::mpdx::GuardFast, F, I, S, Z, 32, false)                  \
  X(float, ::mpdx::GuardFast, F, I, S, Z, 16, false)                  \
  X(float, ::mpdx::GuardFast, F, I, S, Z, 32, true)                   \
  X(float, ::mpdx::GuardFast, F, I, S, Z, 16, true)                   \
  MPDX_IV_24_##I##_##Z(X, F, S)

#define MPDX_IV_S(X, F, I, Z) MPDX_IV_T(X, F, I, true, Z) MPDX_IV_T(X, F, I, false, Z)

#define MPDX_IV_Z(X, F, I) MPDX_IV_S(X, F, I, true) MPDX_IV_S(X, F, I, false)

#define MPDX_INTERIOR_F32_FAST(X)                                     \
  MPDX_IV_Z(X, true, false)                                           \
  MPDX_IV_Z(X, true, true)                                            \
  MPDX_IV_Z(X, false, false)                                          \
  MPDX_IV_Z(X, false, true)

// the mixed-variant interior tasks of ensembles (InteriorMixTask): FP32
// default guards, both tile heights of batched launches, both x / y bound
// treatments.  M(R, G, TY, WXY) is applied to each.
#define MPDX_INTERIOR_MIX_F32_FAST(M)                                 \
  M(float, ::mpdx::GuardFast, 32, false)                              \
  M(float, ::mpdx::GuardFast, 32, true)                               \
  M(float, ::mpdx::GuardFast, 16, false)                              \
  M(float, ::mpdx::GuardFast, 16, true)

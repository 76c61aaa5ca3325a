// Fused MPDATA step kernels (placeholder until the fused path lands).
#pragma once

#include "mpdx_solver.hpp"

namespace mpdx {

inline bool fused_supported(const mpdx_params &, const Layout &) { return false; }
inline int fused_alloc(mpdx_solver *) { return 0; }
inline void fused_free(mpdx_solver *) {}
template <class R, class G> int fused_step(mpdx_solver *) { return MPDX_EUNSUPPORTED; }

}  // namespace mpdx

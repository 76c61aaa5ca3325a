// Batched launches for ensembles of independent solvers.
//
// A task is the argument set of one launch of a fused-path kernel (its
// pointers, layout, box and scales) together with the shape of the grid that
// launch would have (nx x ny x nz blocks).  k_task_list runs many tasks of
// one kernel as a single 1D grid: block b belongs to task q when
// first[q] <= b < first[q + 1] (a binary search over first[], which lives in
// device memory next to the tasks), and runs block b - first[q] of that task
// with the same device code the task's own launch would run.  So a batched
// step computes exactly what each member's own step computes.
#pragma once

#include "mpdx_fused.hpp"
#include "mpdx_pass2d.hpp"

namespace mpdx {

struct TaskBlk {
  unsigned x, y, z;
};

// waves per SIMD a batched launch is compiled for: the interior tasks'
// own (interior_waves_per_eu), 1 (the launch bounds decide) for the others
template <class T> struct TaskWaves {
  static constexpr int v = 1;
};

// the interior tasks' batched launches carry their kernel's occupancy
// attribute in the translation unit that instantiates them
#ifdef MPDX_INTERIOR_TU
#define MPDX_TASK_ATTR __attribute__((amdgpu_waves_per_eu(TaskWaves<T>::v)))
#else
#define MPDX_TASK_ATTR
#endif

// blocks of a task's own grid; first = prefix sums of blocks, n tasks
template <class T>
__global__ void __launch_bounds__(T::kThreads) MPDX_TASK_ATTR k_task_list(const T *tasks, const unsigned *first, int n)
{
  const unsigned b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  // a private copy: the task's fields stay in registers instead of being
  // reloaded from memory after every store the body makes
  const T t = tasks[lo];
  const unsigned l = b - first[lo];
  t.run(TaskBlk{l % t.nx, (l / t.nx) % t.ny, l / (t.nx * t.ny)});
}

// fields every task carries: its grid shape; blocks are kThreads threads,
// kBlockY rows of kThreads / kBlockY (all 1D but the interior kernel's)
struct TaskGrid {
  static constexpr int kBlockY = 1;
  unsigned nx = 1, ny = 1, nz = 1;
  unsigned blocks() const { return nx * ny * nz; }
};

// one box of the scalar halo (k_fz_halo)
template <class R> struct HaloTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *a;
  Layout L;
  BcMap m;
  Box bx;
  __device__ void run(TaskBlk b) const
  {
    int i, j, k;
    if (box_idx_at(bx, b.x, b.y, i, j, k)) fz_halo_cell(a, L, m, i, j, k);
  }
};

// one box of the donor-cell pass (k_fz_donor)
template <class R, bool NPA> struct DonorTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *out;
  const R *in, *C0, *C1, *C2;
  Layout L;
  R in_scale, out_scale;
  int flux_is_C, khn;
  Box bx;
  __device__ void run(TaskBlk b) const
  {
    int i, j, k;
    if (box_idx_at(bx, b.x, b.y, i, j, k))
      fz_donor_cell<R, NPA>(out, in, C0, C1, C2, L, in_scale, out_scale, flux_is_C, khn, i, j, k);
  }
};

// one box of the antidiffusive velocity (k_fz_antidiff)
template <class R, class G, int ND, bool IGA, bool ABS, bool PFC> struct AntidiffTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *V0, *V1, *V2;
  const R *psi, *U0, *U1, *U2;
  Layout L;
  Box bx;
  __device__ void run(TaskBlk b) const
  {
    int i, j, k;
    if (box_idx_at(bx, b.x, b.y, i, j, k)) fz_antidiff_cell<R, G, ND, IGA, ABS, PFC>(V0, V1, V2, psi, U0, U1, U2, L, i, j, k);
  }
};

// one box of the FCT limiter with the donor-cell update (k_fz_fct)
template <class R, class G, int ND, int SHAPE, bool NPA, bool ABS, bool IGA> struct FctTask : TaskGrid {
  static constexpr int kThreads = fz_brick_threads<ND, SHAPE>();
  R *out;
  const R *psi, *psin, *V0, *V1, *V2;
  Layout L;
  R in_scale, out_scale;
  int khn;
  Box bx;
  __device__ void run(TaskBlk b) const
  {
    fz_fct_block<R, G, ND, SHAPE, NPA, ABS, IGA>(out, psi, psin, V0, V1, V2, L, in_scale, out_scale, khn, bx, b.x,
                                                 b.y, b.z);
  }
};

// one halo fill of a scalar or tangential advector component (k_fill_batch)
template <class R> struct FillTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *a;
  Rot r;
  Box bx;
  int side, mode, N, ord;
  __device__ void run(TaskBlk b) const
  {
    int bb, c, m;
    if (ord == 0) {
      if (!box_idx_at(bx, b.x, b.y, bb, m, c)) return;
    } else if (!box_idx_at(bx, b.x, b.y, m, c, bb)) {
      return;
    }
    fill_one(a, r, bb, c, m, side, mode, N);
  }
};

// one side of an open boundary's perpendicular-component zeroing (k_alng_zero)
template <class R> struct AlngZeroTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *av1, *av2;
  Rot r;
  int nd, d, e, raf, ral, rbf, rbl;
  __device__ void run(TaskBlk b) const
  {
    alng_zero_one(av1, av2, r, nd, d, e, raf, ral, rbf, rbl, (int)(b.x * 256u + threadIdx.x));
  }
};

// one box of a normal-component halo fill (k_alng2, k_alng_list)
template <class R> struct AlngTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *avd;
  const R *av1, *av2;
  Rot r;
  Box bx;
  int nd, side, cyc, N, cfirst;
  __device__ void run(TaskBlk b) const
  {
    int ii, bb, c;
    if (box_idx_at(bx, b.x, b.y, ii, bb, c)) alng_one(avd, av1, av2, r, ii, bb, c, nd, side, cyc, N, cfirst);
  }
};

// z halo of the velocity under cyclic z bounds in a ring of rows: x = x0
// and x = x1 with y in [y0, y1], and y = y0 and y = y1 with x in (x0, x1).
// Tangential components: halo cells -m and n-1+m copy cells n-1-m and m;
// the normal one: halo faces -m and n-2+m copy faces n-1-m and m-1
// (bcond/cyclic_3d.hpp:44-58,135-149), m = 1..hw.  One thread per row,
// component, side and m.
template <class R>
__device__ inline void ring_zfill_one(R *v0, R *v1, R *v2, const Layout &L, int x0, int x1, int y0, int y1, int t)
{
  const int hw = L.hw, items = 6 * hw;
  const int nrx = y1 - y0 + 1, nry = x1 - x0 - 1;
  const int r = t / items, k = t - r * items;
  if (r >= 2 * (nrx + nry)) return;
  int x, y;
  if (r < 2 * nrx) {
    x = r < nrx ? x0 : x1;
    y = y0 + r % nrx;
  } else {
    const int q = r - 2 * nrx;
    y = q < nry ? y0 : y1;
    x = x0 + 1 + q % nry;
  }
  const int c = k / (2 * hw), side = (k / hw) % 2, m = 1 + k % hw;
  R *a = c == 0 ? v0 : (c == 1 ? v1 : v2);
  const int n2 = L.n[2];
  const long row = L.org + (long)x * L.st[0] + (long)y * L.st[1];
  int dst, src;
  if (side == 0) { dst = -m; src = n2 - 1 - m; }
  else if (c < 2) { dst = n2 - 1 + m; src = m; }
  else { dst = n2 - 2 + m; src = m - 1; }
  a[row + (long)dst * L.st[2]] = a[row + (long)src * L.st[2]];
}

template <class R> struct RingFillTask : TaskGrid {
  static constexpr int kThreads = 256;
  R *v0, *v1, *v2;
  Layout L;
  int x0, x1, y0, y1;
  long threads() const { return 2L * ((y1 - y0 + 1) + (x1 - x0 - 1)) * 6 * L.hw; }
  __device__ void run(TaskBlk b) const { ring_zfill_one(v0, v1, v2, L, x0, x1, y0, y1, (int)(b.x * 256u + threadIdx.x)); }
};

// one task launched on its own
template <class T>
__global__ void __launch_bounds__(T::kThreads) k_task(const T t)
{
  t.run(TaskBlk{blockIdx.x, blockIdx.y, blockIdx.z});
}

// one launch of the 2D corrective pass (k_fz_pass2d)
template <class R, class G, bool FCT, bool IGA>
struct Pass2dTask : TaskGrid {
  static constexpr int kThreads = kP2Threads;
  Pass2dArgs<R> a;
  __device__ void run(TaskBlk b) const { fz_pass2d_block<R, G, FCT, IGA>(a, b.x); }
};

// a 2D pass of any option variant (variant = FCT | IGA << 1): one batched
// launch holds the pass of every 2D member whatever its options (the
// variants share one set of LDS planes)
template <class R, class G> struct Pass2dMixTask : TaskGrid {
  static constexpr int kThreads = kP2Threads;
  Pass2dArgs<R> a;
  int variant;
  __device__ __forceinline__ void run(TaskBlk b) const
  {
    switch (variant) {
      case 0: fz_pass2d_block<R, G, false, false>(a, b.x); break;
      case 1: fz_pass2d_block<R, G, true, false>(a, b.x); break;
      case 2: fz_pass2d_block<R, G, false, true>(a, b.x); break;
      case 3: fz_pass2d_block<R, G, true, true>(a, b.x); break;
      default: break;
    }
  }
};

// one launch of the interior corrective pass (k_fz_interior); its block
// count is a multiple of 8, so a batch of these keeps the XCD tile order
template <class R, class G, bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV, bool WXY>
struct InteriorTask : TaskGrid {
  static constexpr int kThreads = ITile<TYV>::Y * ITile<TYV>::Z, kBlockY = ITile<TYV>::Y;
  InteriorArgs<R> a;
  __device__ void run(TaskBlk b) const { fz_interior_block<R, G, FCT, IGA, STOREV, FUSE0, TYV, WXY>(a, b.x); }
};
template <class R, class G, bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV, bool WXY>
struct TaskWaves<InteriorTask<R, G, FCT, IGA, STOREV, FUSE0, TYV, WXY>> {
  static constexpr int v = interior_waves_per_eu<R, TYV>();
};

// An interior pass of any option variant (variant = FCT | IGA << 1 |
// STOREV << 2 | FUSE0 << 3) with one element type, guard, tile height and
// x / y treatment, so that one batched launch holds the interior passes of
// every member whatever options and pass each is at.  The variants share
// one LDS array of the largest size.  iga takes only its single corrective
// pass (FUSE0 without STOREV; iga allows n_iters <= 2), so ten variants.
constexpr bool interior_mix_variant(bool fct, bool iga, bool storev, bool fuse0)
{
  (void)fct;
  return !iga || (fuse0 && !storev);
}
template <class R, class G, int TYV, bool WXY> struct InteriorMixTask : TaskGrid {
  static constexpr int kThreads = ITile<TYV>::Y * ITile<TYV>::Z, kBlockY = ITile<TYV>::Y;
  static constexpr int cmax(int x, int y) { return x > y ? x : y; }
  static constexpr int kLds = cmax(cmax(interior_lds_elems<true, true, TYV>(), interior_lds_elems<true, false, TYV>()),
                                   cmax(interior_lds_elems<false, true, TYV>(), interior_lds_elems<false, false, TYV>()));
  InteriorArgs<R> a;
  int variant;
  __device__ __forceinline__ void run(TaskBlk b) const
  {
    switch (variant) {
#define MPDX_MIX(F, I, S, Z)                                                                     \
  case (F) | (I) << 1 | (S) << 2 | (Z) << 3:                                                     \
    fz_interior_block<R, G, F, I, S, Z, TYV, WXY, kLds>(a, b.x);                                 \
    break;
      MPDX_MIX(1, 0, 1, 1) MPDX_MIX(1, 0, 1, 0) MPDX_MIX(1, 0, 0, 1) MPDX_MIX(1, 0, 0, 0)
      MPDX_MIX(0, 0, 1, 1) MPDX_MIX(0, 0, 1, 0) MPDX_MIX(0, 0, 0, 1) MPDX_MIX(0, 0, 0, 0)
      MPDX_MIX(1, 1, 0, 1) MPDX_MIX(0, 1, 0, 1)
#undef MPDX_MIX
      default: break;
    }
  }
};
template <class R, class G, int TYV, bool WXY> struct TaskWaves<InteriorMixTask<R, G, TYV, WXY>> {
  static constexpr int v = interior_waves_per_eu<R, TYV>();
};

}  // namespace mpdx

// Shared device helpers for the MPDATA kernels (gfx950).
//
// Arrays live in HBM as halo-extended boxes: dimension d spans
// [-hw, n_d-1+hw] (scalars) or the face range [-hw, n_d-2+hw] of a
// staggered component (face index i <-> i+1/2, formulae/arakawa_c.hpp:17-41).
// Rows of the innermost dimension start on 128-byte boundaries, with index 0
// placed so that the boundary strips [-2, 4] and [n-5, n+1] each fall in one
// 128-byte line where possible.  Element (i,j,k) is at
// org + i*st[0] + j*st[1] + k*st[2]; in 2D st[2] == 0.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <utility>
#include <vector>

#include "mpdx_fastdiv.hpp"

namespace mpdx {

enum : unsigned {
  O_FCT = 1u << 0, O_ABS = 1u << 1, O_TOT = 1u << 2, O_PFC = 1u << 3,
  O_NPA = 1u << 4, O_IGA = 1u << 5, O_NUG = 1u << 6, O_DFL = 1u << 7,
  O_KHN = 1u << 8
};
enum { BC_CYCLIC = 0, BC_OPEN = 1 };

struct Layout {
  int nd, hw;
  int n[3];
  long st[3];
  long org;
  long size;      // elements per array, including padding
};

// index rotation idxperm::pi<d> (formulae/idxperm.hpp:96-125):
// position[d] = a, position[d+1] = b, position[d+2] = c
struct Rot {
  long o, sA, sB, sC;
};

inline Rot make_rot(const Layout &L, int d)
{
  Rot r;
  r.o = L.org;
  r.sA = L.st[d];
  r.sB = L.st[(d + 1) % L.nd];
  r.sC = L.nd == 3 ? L.st[(d + 2) % 3] : 0;
  return r;
}

__host__ __device__ inline long rr(const Rot &r, int a, int b, int c)
{
  return r.o + (long)a * r.sA + (long)b * r.sB + (long)c * r.sC;
}

// index box (a, b, c), c fastest.  Box kernels run on a 2D grid: blockIdx.y
// selects a, blockIdx.x * blockDim.x + threadIdx.x the (b, c) pair within
// the plane, split by one 32-bit multiply-high.
struct Box {
  int lo[3];
  int ext[3];
  long total;
  unsigned plane;   // ext[1] * ext[2]
  FastDiv d1, d2;   // division by ext[1], ext[2]
};

inline Box make_box(int a0, int a1, int b0, int b1, int c0, int c1)
{
  Box b;
  b.lo[0] = a0; b.lo[1] = b0; b.lo[2] = c0;
  b.ext[0] = a1 - a0 + 1; b.ext[1] = b1 - b0 + 1; b.ext[2] = c1 - c0 + 1;
  if (b.ext[0] <= 0 || b.ext[1] <= 0 || b.ext[2] <= 0) {
    b.total = 0;
    b.plane = 0;
    b.d1 = make_fastdiv(1);
    b.d2 = make_fastdiv(1);
  } else {
    b.total = (long)b.ext[0] * b.ext[1] * b.ext[2];
    b.plane = (unsigned)((long)b.ext[1] * b.ext[2]);
    b.d1 = make_fastdiv((unsigned)b.ext[1]);
    b.d2 = make_fastdiv((unsigned)b.ext[2]);
  }
  return b;
}

// Up to 16 boxes of different shapes launched as one 1D grid of 256-thread
// blocks: block k belongs to box q when first[q] <= k < first[q + 1], and a
// box's cells are numbered (a, b, c), c fastest.  tag is free for the
// kernel (e.g. the side of a halo).
struct BoxList {
  Box bx[16];
  unsigned first[17];
  int tag[16];
  int n;
};

inline void boxlist_add(BoxList &l, const Box &b, int tag = 0)
{
  if (b.total <= 0) return;
  if (l.n == 0) l.first[0] = 0;
  l.bx[l.n] = b;
  l.tag[l.n] = tag;
  l.first[l.n + 1] = l.first[l.n] + (unsigned)((b.total + 255) / 256);
  ++l.n;
}

inline unsigned boxlist_blocks(const BoxList &l) { return l.n ? l.first[l.n] : 0; }

// Launches (box, tag) pairs as box lists: a box of 2^31 cells or more is cut
// along its first coordinate (the flat index within a box is 32-bit), and
// the pieces go out in lists of at most 16 boxes and 2^31 blocks.
// launch(const BoxList &) issues one kernel.
template <class F>
inline void boxlist_launch(const std::vector<std::pair<Box, int>> &boxes, F launch)
{
  BoxList l{};
  auto flush = [&]() {
    if (l.n) launch(l);
    l = BoxList{};
  };
  for (const auto &bt : boxes) {
    const Box &b = bt.first;
    if (b.total <= 0) continue;
    const long plane = (long)b.ext[1] * b.ext[2];
    const long rows = std::max(1L, (long)((1L << 31) - 1) / plane);
    for (long a0 = 0; a0 < b.ext[0]; a0 += rows) {
      const int a1 = (int)std::min<long>(b.ext[0] - 1, a0 + rows - 1);
      const Box p = make_box(b.lo[0] + (int)a0, b.lo[0] + a1, b.lo[1], b.lo[1] + b.ext[1] - 1, b.lo[2],
                             b.lo[2] + b.ext[2] - 1);
      if (l.n == 16 || (long)boxlist_blocks(l) + (p.total + 255) / 256 >= (1L << 31)) flush();
      boxlist_add(l, p, bt.second);
    }
  }
  flush();
}

__device__ inline bool boxlist_idx(const BoxList &l, int &q, int &a, int &bb, int &c)
{
  const unsigned blk = blockIdx.x;
  int s = 0;
  for (int i = 1; i < l.n; ++i) s += blk >= l.first[i];
  q = s;
  const Box &b = l.bx[s];
  const unsigned t = (blk - l.first[s]) * blockDim.x + threadIdx.x;
  if ((long)t >= b.total) return false;
  const unsigned r = fdiv(t, b.d2);   // a * ext[1] + b
  c = (int)(t - r * (unsigned)b.ext[2]) + b.lo[2];
  const unsigned aa = fdiv(r, b.d1);
  bb = (int)(r - aa * (unsigned)b.ext[1]) + b.lo[1];
  a = (int)aa + b.lo[0];
  return true;
}

__device__ inline bool box_idx(const Box &b, int &a, int &bb, int &c)
{
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.plane) return false;
  const unsigned q = fdiv(t, b.d2);
  a = (int)blockIdx.y + b.lo[0];
  bb = (int)q + b.lo[1];
  c = (int)(t - q * (unsigned)b.ext[2]) + b.lo[2];
  return true;
}

template <class G> struct lim;
template <> struct lim<float> {
  __host__ __device__ static constexpr float tiny() { return FLT_MIN; }
  __host__ __device__ static constexpr float eps() { return FLT_EPSILON; }
};
template <> struct lim<double> {
  __host__ __device__ static constexpr double tiny() { return DBL_MIN; }
  __host__ __device__ static constexpr double eps() { return DBL_EPSILON; }
};

// std::max / std::min argument order (ties and NaNs resolve like the reference)
template <class T> __host__ __device__ inline T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> __host__ __device__ inline T smin(T a, T b) { return (b < a) ? b : a; }

// pospart / negpart: formulae/common.hpp:87-124
template <bool NPA, class R> __device__ inline R pospart(R x)
{
  if (NPA) return (x + fabs(x)) / 2;
  return smax(R(0), x);
}
template <bool NPA, class R> __device__ inline R negpart(R x)
{
  if (NPA) return (x - fabs(x)) / 2;
  return smin(R(0), x);
}

// Guard policy for FP32 state: the FP64 guard constants (epsilon<double>
// for fct_frac; the smallest normal float stands in for tiny<double>, which
// underflows in FP32) with the ratio formed by the hardware reciprocal
// (v_rcp_f32, <= 1 ulp) times the numerator.
struct GuardFast {};

// frac / fct_frac: formulae/mpdata/formulae_mpdata_common.hpp:61-98, with the
// guard evaluated in precision G (G = real_t reproduces the reference)
template <bool PFC, class G, class R> __device__ inline R frac(R nom, R den)
{
  if (PFC) return den != 0 ? R(G(nom) / G(den)) : R(0);
  return R(G(nom) / (G(den) + lim<G>::tiny()));
}
template <class G, class R> __device__ inline R fct_frac(R nom, R den)
{
  return R(G(nom) / (G(den) + lim<G>::eps()));
}
template <> __device__ inline float frac<true, GuardFast, float>(float nom, float den)
{
  return den != 0 ? nom * __builtin_amdgcn_rcpf(den) : 0.f;
}
template <> __device__ inline float frac<false, GuardFast, float>(float nom, float den)
{
  return nom * __builtin_amdgcn_rcpf(den + FLT_MIN);
}
template <> __device__ inline float fct_frac<GuardFast, float>(float nom, float den)
{
  return nom * __builtin_amdgcn_rcpf(den + float(DBL_EPSILON));
}

// donor-cell flux: formulae/donorcell_formulae.hpp:25-34
template <bool NPA, class R> __device__ inline R donor_flux(R psi_l, R psi_r, R C)
{
  return pospart<NPA>(C) * psi_l + negpart<NPA>(C) * psi_r;
}

}  // namespace mpdx

// Global reductions over the interior of a field: min, max, the FP64
// compensated sum (sharedmem::sum with sum_khn, sharedmem.hpp:115-155,
// kahan_reduction.hpp:15-58), and the two advector diagnostics that reduce
// a per-cell value to its maximum: |div GC| (solver_3d.hpp:400-409,
// solver_2d.hpp:305-313) and the Courant number (solver_3d.hpp:390-398,
// solver_2d.hpp:295-303), computed on the fly rather than stored.
//
// Two launches.  k_red_part: block b covers a contiguous run of rows of one
// x plane (plane b / S, part b % S of it, S parts per plane); each thread
// keeps an FP64 accumulator over its cells of each row, the 64 lanes of a wave
// combine theirs with cross-lane shuffles (a butterfly), the waves of the
// block through LDS, and the block writes one partial.  k_red_final: one
// block combines the partials the same way.  Sums are compensated at every
// level: each accumulator is a pair (hi, lo) with value hi + lo, an addition
// adds the exact rounding error of hi + x (TwoSum, Knuth TAOCP vol. 2,
// 4.2.2) into lo, and two pairs merge as TwoSum of the his plus the sum of
// the los.  The order of every operation is fixed by the launch shape, so a
// result is the same in every run.
#pragma once

#include "mpdx_common.hpp"

namespace mpdx {

enum RedKind : int { RK_MIN = 0, RK_MAX = 1, RK_SUM = 2, RK_DIV = 3, RK_CFL = 4 };

constexpr int kRedThreads = 256;      // threads of a k_red_part block
constexpr int kRedFinalThreads = 1024;
constexpr int kRedParts = 4096;       // partials aimed at (S = kRedParts / planes, at least 1)

// parts per x plane and partials of a reduction over n0 planes
__host__ __device__ inline int red_split(int n0) { return n0 >= kRedParts ? 1 : (kRedParts / n0 > 64 ? 64 : kRedParts / n0); }
__host__ __device__ inline int red_parts(int n0) { return n0 * red_split(n0); }

__device__ __forceinline__ void acc_add(double &hi, double &lo, double x)
{
  const double s = hi + x;
  const double v = s - hi;
  lo += (hi - (s - v)) + (x - v);
  hi = s;
}

__device__ __forceinline__ void acc_merge(double &hi, double &lo, double hi2, double lo2)
{
  const double s = hi + hi2;
  const double v = s - hi;
  const double err = (hi - (s - v)) + (hi2 - v);
  lo = (lo + lo2) + err;
  hi = s;
}

template <int K> __device__ __forceinline__ double red_identity()
{
  return K == RK_MIN ? INFINITY : (K == RK_SUM ? 0.0 : -INFINITY);
}

template <int K> __device__ __forceinline__ void red_combine(double &hi, double &lo, double hi2, double lo2)
{
  if constexpr (K == RK_SUM) acc_merge(hi, lo, hi2, lo2);
  else if constexpr (K == RK_MIN) hi = smin(hi, hi2);
  else hi = smax(hi, hi2);
}

// a block's (hi, lo) over all its threads, valid in thread 0: wave
// butterflies, then wave 0 over the per-wave results
template <int K, int NT>
__device__ inline void red_block(double &hi, double &lo)
{
  constexpr int NW = NT / 64;
  __shared__ double sh_hi[NW], sh_lo[NW];
  for (int o = 32; o >= 1; o >>= 1) {
    const double h2 = __shfl_xor(hi, o), l2 = __shfl_xor(lo, o);
    red_combine<K>(hi, lo, h2, l2);
  }
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) { sh_hi[w] = hi; sh_lo[w] = lo; }
  __syncthreads();
  if (w == 0) {
    hi = lane < NW ? sh_hi[lane] : red_identity<K>();
    lo = lane < NW ? sh_lo[lane] : 0.0;
    for (int o = 32; o >= 1; o >>= 1) {
      const double h2 = __shfl_xor(hi, o), l2 = __shfl_xor(lo, o);
      red_combine<K>(hi, lo, h2, l2);
    }
  }
}

// the per-cell value of reduction kind K at element p
template <int K, class R>
__device__ __forceinline__ double red_value(const R *a, const R *g1, const R *g2, const Layout &L, long p)
{
  if constexpr (K == RK_DIV) {
    R v;
    if (L.nd == 3)
      v = fabs((a[p] - a[p - L.st[0]]) + (g1[p] - g1[p - L.st[1]]) + (g2[p] - g2[p - L.st[2]])) / 1;
    else
      v = fabs((a[p] - a[p - L.st[0]]) + (g1[p] - g1[p - L.st[1]])) / 1;
    return double(v);
  } else if constexpr (K == RK_CFL) {
    R v;
    if (L.nd == 3)
      v = R(0.5) * (fabs(a[p] + a[p - L.st[0]]) + fabs(g1[p] + g1[p - L.st[1]]) + fabs(g2[p] + g2[p - L.st[2]])) / 1;
    else
      v = R(0.5) * (fabs(a[p] + a[p - L.st[0]]) + fabs(g1[p] + g1[p - L.st[1]])) / 1;
    return double(v);
  } else {
    return double(a[p]);
  }
}

// partial b of kind K over the box bx (bx.ext[0] planes, S parts each);
// RK_DIV / RK_CFL read the advector components a, g1, g2.  A plane is walked
// as rows of contiguous cells (3D: ext[1] rows of ext[2] cells along z; 2D:
// one row of ext[1] cells along y); the S parts of a plane split its rows,
// or, with a single row, the row itself.
template <int K, class R>
__global__ void __launch_bounds__(kRedThreads)
k_red_part(double2 *part, const R *a, const R *g1, const R *g2, Layout L, Box bx, int S)
{
  const int b = blockIdx.x;
  const int plane_i = b / S, s = b - plane_i * S;
  const bool d3 = L.nd == 3;
  const int rows = d3 ? bx.ext[1] : 1, len = d3 ? bx.ext[2] : bx.ext[1];
  int r0 = 0, r1 = rows, k0 = 0, k1 = len;
  if (rows > 1) {
    r0 = (int)((long)rows * s / S);
    r1 = (int)((long)rows * (s + 1) / S);
  } else {
    k0 = (int)((long)len * s / S);
    k1 = (int)((long)len * (s + 1) / S);
  }
  // element of (row r, cell k): first + r * rst + k
  const long first = L.org + (long)(plane_i + bx.lo[0]) * L.st[0] + (long)bx.lo[1] * L.st[1] +
                     (d3 ? (long)bx.lo[2] * L.st[2] : 0);
  const long rst = d3 ? L.st[1] : 0;
  double hi = red_identity<K>(), lo = 0.0;
  for (int r = r0; r < r1; ++r) {
    const long row = first + (long)r * rst;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 2 * kRedThreads) {
      const bool two = k + kRedThreads < k1;
      const double v0 = red_value<K>(a, g1, g2, L, row + k);
      const double v1 = two ? red_value<K>(a, g1, g2, L, row + k + kRedThreads) : red_identity<K>();
      if constexpr (K == RK_MIN) hi = smin(smin(hi, v0), v1);
      else if constexpr (K == RK_MAX || K == RK_DIV || K == RK_CFL) hi = smax(smax(hi, v0), v1);
      else { acc_add(hi, lo, v0); acc_add(hi, lo, v1); }
    }
  }
  constexpr int KC = K == RK_SUM ? RK_SUM : (K == RK_MIN ? RK_MIN : RK_MAX);
  red_block<KC, kRedThreads>(hi, lo);
  if (threadIdx.x == 0) part[b] = make_double2(hi, lo);
}

// the n partials combined into *out (hi + lo for a sum)
template <int K>
__global__ void __launch_bounds__(kRedFinalThreads) k_red_final(double *out, const double2 *part, int n)
{
  constexpr int KC = K == RK_SUM ? RK_SUM : (K == RK_MIN ? RK_MIN : RK_MAX);
  double hi = red_identity<KC>(), lo = 0.0;
  for (int q = threadIdx.x; q < n; q += kRedFinalThreads) red_combine<KC>(hi, lo, part[q].x, part[q].y);
  red_block<KC, kRedFinalThreads>(hi, lo);
  if (threadIdx.x == 0) *out = KC == RK_SUM ? hi + lo : hi;
}

}  // namespace mpdx

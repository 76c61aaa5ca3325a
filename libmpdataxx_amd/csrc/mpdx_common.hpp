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

#include <cfloat>
#include <cstdint>

#include "mpdx_box.hpp"
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

// the cell of a box-list launch's thread (blocks of 256 threads)
__device__ inline bool boxlist_idx(const BoxList &l, int &q, int &a, int &bb, int &c)
{
  return boxlist_cell(l, blockIdx.x, threadIdx.x, q, a, bb, c);
}

// cell of this thread in block (gx, gy) of a Box launch of 256-thread blocks
__device__ inline bool box_idx_at(const Box &b, unsigned gx, unsigned gy, int &a, int &bb, int &c)
{
  const unsigned t = gx * 256u + threadIdx.x;
  if (t >= b.plane) return false;
  const unsigned q = fdiv(t, b.d2);
  a = (int)gy + b.lo[0];
  bb = (int)q + b.lo[1];
  c = (int)(t - q * (unsigned)b.ext[2]) + b.lo[2];
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

// FP64 quotient of the guarded ratios: the hardware reciprocal (v_rcp_f64,
// relative error below 2^-24 on gfx950: 2^-24.4 at most over 4M random
// operands) refined by one Newton-Raphson step r <- r + r (1 - d r) to about
// 2^-48, the quotient q = n r, and one correction q + r (n - d q) whose
// residual the fused multiply-add forms exactly.  Within 1 ulp of the IEEE
// quotient (equal to it for each of those 4M operands) for DBL_MIN <= |d|
// and a finite quotient, in 6 VALU instructions where the IEEE sequence
// (div_scale, div_fmas, div_fixup) takes 11 and hazard waits.  A
// denominator that cancels to 0 or to a subnormal (a guarded sum within
// 2 DBL_MIN of zero) gives NaN where IEEE division gives an infinity.
__device__ __forceinline__ double ddiv(double n, double d)
{
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}
template <class G> __device__ __forceinline__ G gdiv(G n, G d) { return n / d; }
template <> __device__ __forceinline__ double gdiv<double>(double n, double d) { return ddiv(n, d); }

// frac / fct_frac: formulae/mpdata/formulae_mpdata_common.hpp:61-98, with the
// guard evaluated in precision G (G = real_t reproduces the reference's
// constants; FP32 divides as IEEE does, FP64 through ddiv)
template <bool PFC, class G, class R> __device__ inline R frac(R nom, R den)
{
  if (PFC) return den != 0 ? R(gdiv<G>(G(nom), G(den))) : R(0);
  return R(gdiv<G>(G(nom), G(den) + lim<G>::tiny()));
}
template <class G, class R> __device__ inline R fct_frac(R nom, R den)
{
  return R(gdiv<G>(G(nom), G(den) + lim<G>::eps()));
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

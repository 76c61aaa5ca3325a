// Division by a run-time invariant 32-bit divisor, shared by host and
// device code (the host half is what the CPU test of it compiles).
#pragma once

#ifndef __HIPCC__
#define MPDX_HD
#else
#define MPDX_HD __host__ __device__
#endif

namespace mpdx {

// unsigned 32-bit division by a run-time invariant divisor d >= 1 for
// dividends n < 2^31 (round-up multiplier, Granlund & Montgomery 1994,
// "Division by invariant integers using multiplication", sec. 4):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, q = (mulhi(n, m) + n) >> l
struct FastDiv {
  unsigned m;
  unsigned l;
};

inline FastDiv make_fastdiv(unsigned d)
{
  FastDiv f;
  f.l = 0;
  while ((1ull << f.l) < d) ++f.l;
  f.m = (unsigned)((((1ull << f.l) - d) << 32) / d + 1);
  return f;
}

MPDX_HD inline unsigned fdiv(unsigned n, FastDiv f)
{
#ifdef __HIP_DEVICE_COMPILE__
  const unsigned t = __umulhi(n, f.m);
#else
  const unsigned t = (unsigned)(((unsigned long long)n * f.m) >> 32);
#endif
  return (t + n) >> f.l;
}

}  // namespace mpdx

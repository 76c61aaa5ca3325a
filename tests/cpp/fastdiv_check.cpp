// Host check of the multiply-high division used by the box kernels: for a
// spread of divisors (every d <= 4096, powers of two and their neighbours,
// large odd values) and dividends (0..8191, near multiples of d, up to
// 2^31 - 1) fdiv(n, make_fastdiv(d)) must equal n / d.
#include "mpdx_fastdiv.hpp"

#include <cstdio>
#include <vector>

int main()
{
  std::vector<unsigned> ds;
  for (unsigned d = 1; d <= 4096; ++d) ds.push_back(d);
  for (int k = 12; k < 31; ++k) {
    ds.push_back((1u << k) - 1);
    ds.push_back(1u << k);
    ds.push_back((1u << k) + 1);
  }
  for (unsigned d : {46337u, 65535u, 262147u, 1000003u, 2147483647u}) ds.push_back(d);
  long bad = 0, checked = 0;
  for (unsigned d : ds) {
    const mpdx::FastDiv f = mpdx::make_fastdiv(d);
    auto chk = [&](unsigned long long n) {
      if (n >= (1ull << 31)) return;
      ++checked;
      if (mpdx::fdiv((unsigned)n, f) != (unsigned)(n / d)) {
        if (bad++ < 10) std::printf("mismatch d=%u n=%llu\n", d, n);
      }
    };
    for (unsigned n = 0; n < 8192; ++n) chk(n);
    for (unsigned long long q = 1; q < (1ull << 31) / d; q = q * 3 + 1)
      for (int e = -2; e <= 2; ++e) chk(q * d + e);
    for (int e = 1; e <= 64; ++e) chk((1ull << 31) - e);
  }
  std::printf("fastdiv: %ld dividends checked, %ld mismatches\n", checked, bad);
  return bad != 0;
}

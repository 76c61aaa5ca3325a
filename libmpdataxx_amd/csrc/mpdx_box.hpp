// Index boxes and box lists, shared by host and device code (the host half
// is what the CPU test of it compiles).
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

#include "mpdx_fastdiv.hpp"

namespace mpdx {

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

// cell (a, b, c) of box q for thread thr of block blk of a box-list launch
// (256-thread blocks); false for the unused threads of a box's last block
MPDX_HD inline bool boxlist_cell(const BoxList &l, unsigned blk, unsigned thr, int &q, int &a, int &bb, int &c)
{
  int s = 0;
  for (int i = 1; i < l.n; ++i) s += blk >= l.first[i];
  q = s;
  const Box &b = l.bx[s];
  const unsigned t = (blk - l.first[s]) * 256u + thr;
  if ((long)t >= b.total) return false;
  const unsigned r = fdiv(t, b.d2);   // a * ext[1] + b
  c = (int)(t - r * (unsigned)b.ext[2]) + b.lo[2];
  const unsigned aa = fdiv(r, b.d1);
  bb = (int)(r - aa * (unsigned)b.ext[1]) + b.lo[1];
  a = (int)aa + b.lo[0];
  return true;
}

}  // namespace mpdx

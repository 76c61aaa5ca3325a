// Host check of box-list launches (mpdx_box.hpp), the index space of the
// halo kernels: every thread of every block that boxlist_launch issues is
// decoded with boxlist_cell; each cell of each box must come out exactly
// once, with the box's tag, and no thread may decode to a cell outside its
// box.  Lists of up to 20 boxes of mixed shapes (empty ones, negative
// origins, one-cell extents) exercise the block prefix table and the split
// into several launches; boxes of 2^31 cells or more must be cut along their
// first coordinate into pieces below 2^31 cells that tile the box.
#include "mpdx_box.hpp"

#include <cstdio>
#include <map>
#include <random>
#include <tuple>

using mpdx::Box;
using mpdx::BoxList;

static long bad = 0;

static void fail(const char *what)
{
  if (bad++ < 10) std::printf("%s\n", what);
}

// enumerate every launched thread; count each decoded cell per (tag, cell),
// tags being unique per box
static void check_list(const std::vector<std::pair<Box, int>> &boxes)
{
  std::map<std::tuple<int, int, int, int>, int> seen;
  int launches = 0;
  mpdx::boxlist_launch(boxes, [&](const BoxList &l) {
    ++launches;
    if (l.n < 1 || l.n > 16) fail("list size out of range");
    for (unsigned blk = 0; blk < mpdx::boxlist_blocks(l); ++blk)
      for (unsigned thr = 0; thr < 256; ++thr) {
        int q, a, b, c;
        if (!mpdx::boxlist_cell(l, blk, thr, q, a, b, c)) continue;
        const Box &x = l.bx[q];
        if (blk < l.first[q] || blk >= l.first[q + 1]) fail("block outside its box's range");
        if (a < x.lo[0] || a >= x.lo[0] + x.ext[0] || b < x.lo[1] || b >= x.lo[1] + x.ext[1] ||
            c < x.lo[2] || c >= x.lo[2] + x.ext[2])
          fail("cell outside its box");
        ++seen[{l.tag[q], a, b, c}];
      }
  });
  long expected = 0;
  for (const auto &bt : boxes) {
    const Box &x = bt.first;
    if (x.total <= 0) continue;
    expected += x.total;
    for (int a = 0; a < x.ext[0]; ++a)
      for (int b = 0; b < x.ext[1]; ++b)
        for (int c = 0; c < x.ext[2]; ++c) {
          auto it = seen.find({bt.second, x.lo[0] + a, x.lo[1] + b, x.lo[2] + c});
          if (it == seen.end() || it->second != 1) fail("cell missing or decoded twice");
        }
  }
  long got = 0;
  for (const auto &kv : seen) got += kv.second;
  if (got != expected) fail("decoded cell count differs from the boxes' total");
  int nonempty = 0;
  for (const auto &bt : boxes) nonempty += bt.first.total > 0;
  if (nonempty > 16 && launches < 2) fail("more than 16 boxes in one launch");
}

int main()
{
  std::mt19937 rng(12345);
  std::uniform_int_distribution<int> ext(0, 9), org(-4, 20), cnt(1, 20);
  int lists = 0;
  for (int trial = 0; trial < 150; ++trial) {
    std::vector<std::pair<Box, int>> boxes;
    const int n = cnt(rng);
    for (int i = 0; i < n; ++i) {
      const int a0 = org(rng), b0 = org(rng), c0 = org(rng);
      // ext 0 gives an empty box (skipped), 1 a one-cell extent
      const Box x = mpdx::make_box(a0, a0 + ext(rng) - 1, b0, b0 + ext(rng) * 7 - 1, c0, c0 + ext(rng) * 13 - 1);
      boxes.emplace_back(x, i);   // unique tag: random boxes may overlap
    }
    check_list(boxes);
    ++lists;
  }
  // a halo-plane sized box with a many-row first coordinate: cut below 2^31
  // cells per piece, pieces contiguous along a and covering the box
  {
    const Box big = mpdx::make_box(-2, 4097, -2, 1023 + 2, -2, 1023 + 2);   // 4100 x 1028 x 1028
    if (big.total < (1L << 31)) fail("test box too small");
    long covered = 0;
    int next_a = big.lo[0];
    mpdx::boxlist_launch({{big, 7}}, [&](const BoxList &l) {
      for (int q = 0; q < l.n; ++q) {
        const Box &p = l.bx[q];
        if (p.total >= (1L << 31)) fail("piece of 2^31 cells or more");
        if (p.lo[0] != next_a) fail("pieces not contiguous along the first coordinate");
        if (p.lo[1] != big.lo[1] || p.ext[1] != big.ext[1] || p.lo[2] != big.lo[2] || p.ext[2] != big.ext[2])
          fail("piece changes the plane");
        if (l.tag[q] != 7) fail("piece lost its tag");
        next_a = p.lo[0] + p.ext[0];
        covered += p.total;
        // the last thread of the piece decodes to its last cell
        int qq, a, b, c;
        const unsigned blk = l.first[q] + (unsigned)((p.total - 1) / 256), thr = (unsigned)((p.total - 1) % 256);
        if (!mpdx::boxlist_cell(l, blk, thr, qq, a, b, c) || qq != q || a != p.lo[0] + p.ext[0] - 1 ||
            b != p.lo[1] + p.ext[1] - 1 || c != p.lo[2] + p.ext[2] - 1)
          fail("last cell of a piece decoded wrong");
      }
    });
    if (covered != big.total || next_a != big.lo[0] + big.ext[0]) fail("pieces do not tile the box");
  }
  std::printf("%d lists, %ld mismatches\n", lists, bad);
  return bad ? 1 : 0;
}

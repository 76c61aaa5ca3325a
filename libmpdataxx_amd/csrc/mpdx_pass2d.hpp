// 2D corrective pass in one launch.
//
// A block owns a 16 x 16 tile of output cells and stages everything the
// tile's update reads in LDS, over the tile plus a frame of 3 cells: psi^n
// (scaled by hint_scale) and the advector, then psi^k (loaded, or with FUSE0
// computed from psi^n and GC by the donor-cell formula on the tile plus 2
// cells, donorcell_formulae.hpp:50-79), the antidiffusive velocity on the
// faces of the tile plus 1 cell (formulae_mpdata_2d.hpp:241-278), beta_up /
// beta_dn on the tile plus 1 cell (formulae_mpdata_fct_2d.hpp:21-120) and
// finally the limited fluxes and the donor-cell update of the tile's cells
// (one thread per cell).  Each quantity is evaluated with the same device
// functions and operand order as the 2D boundary-shell kernels (fz_donor_cell,
// antidiff_at, beta_at, fz_fct_block).
//
// Loads: interior tiles (lo >= 2, hi <= n-3) read the frame from the
// materialized halos.  wrap (cyclic bounds on both sides of x and y, fields
// repeating plane 0 at n-1 along both): the tiles cover every cell and read
// a frame cell at its periodic image in 0..n-1, the value the cyclic halo
// fill copies there (bcond/cyclic_2d.hpp); no halo is read.
#pragma once

#include "mpdx_fused.hpp"

namespace mpdx {

constexpr int kP2T = 16;            // output tile side (one cell per thread)
constexpr int kP2W = kP2T + 6;      // the staged region: tile plus a 3-cell frame
constexpr int kP2Threads = kP2T * kP2T;

template <class R>
struct Pass2dArgs {
  R *out;
  const R *psi, *psin;   // psi^k (unused with fuse0) and psi^n
  const R *U0, *U1;
  R *V0, *V1;            // unlimited velocity out (storev), may be null
  Layout L;
  int lo[2], hi[2];      // output cells (x, y), inclusive
  int vlo[2], vhi[2];    // cells whose velocity is stored
  int nb[2];             // tiles along y and x
  R in_scale, out_scale;
  int fuse0, storev, wrap;
};

// the LDS planes of a 2D pass block (K: psi^n, psi^k, the two advector
// components, the two velocity components, beta up and down), one set per
// element type, so that a kernel whose blocks run one of several variants
// (Pass2dMixTask) holds one set
template <class R, int K> __shared__ R pass2d_lds[kP2W * kP2W];

template <class R, class G, bool FCT, bool IGA>
__device__ __forceinline__ void fz_pass2d_block(const Pass2dArgs<R> &a, unsigned blk)
{
#define s_pn (pass2d_lds<R, 0>)
#define s_pk (pass2d_lds<R, 1>)
#define s_ux (pass2d_lds<R, 2>)
#define s_uy (pass2d_lds<R, 3>)
#define s_vx (pass2d_lds<R, 4>)
#define s_vy (pass2d_lds<R, 5>)
#define s_bu (pass2d_lds<R, 6>)
#define s_bd (pass2d_lds<R, 7>)
  const Layout &L = a.L;
  const int n0 = L.n[0], n1 = L.n[1];
  const int by = (int)(blk % (unsigned)a.nb[0]), bx = (int)(blk / (unsigned)a.nb[0]);
  const int x0 = a.lo[0] + bx * kP2T, y0 = a.lo[1] + by * kP2T;
  const int tid = threadIdx.x;
  // LDS slot of tile offset (i, j), i, j in -3 .. kP2T + 2
  auto S = [](int i, int j) { return (i + 3) * kP2W + (j + 3); };
  // memory coordinate of a frame cell: its periodic image (wrap; cells past
  // n+2 feed no output and are clamped), or itself within the halo (an
  // interior tile's frame reaches at most one cell past the domain; cells
  // past the last tile's end feed no output and are clamped to that halo
  // cell)
  auto mc = [&](int c, int n) {
    if (a.wrap) return c < 0 ? c + (n - 1) : (c > n - 1 ? min(c - (n - 1), n - 1) : c);
    return min(c, n);
  };
  auto gidx = [&](int i, int j) -> long {
    return L.org + (long)mc(x0 + i, n0) * L.st[0] + (long)mc(y0 + j, n1) * L.st[1];
  };

  // ---- psi^n (scaled), psi^k and the advector over the staged region
  for (int q = tid; q < kP2W * kP2W; q += kP2Threads) {
    const int i = q / kP2W - 3, j = q % kP2W - 3;
    const long p = gidx(i, j);
    if (FCT || a.fuse0) s_pn[q] = a.psin[p] * a.in_scale;
    if (!a.fuse0) s_pk[q] = a.psi[p];
    s_ux[q] = a.U0[p];
    s_uy[q] = a.U1[p];
  }
  __syncthreads();

  // ---- FUSE0: psi^1 = donor cell of psi^n with GC over the tile plus 2
  if (a.fuse0) {
    constexpr int W2 = kP2T + 4;
    for (int q = tid; q < W2 * W2; q += kP2Threads) {
      const int i = q / W2 - 2, j = q % W2 - 2;
      const R pc = s_pn[S(i, j)];
      const R f0 = donor_flux<false>(pc, s_pn[S(i + 1, j)], s_ux[S(i, j)]);
      const R f1 = donor_flux<false>(s_pn[S(i - 1, j)], pc, s_ux[S(i - 1, j)]);
      const R f2 = donor_flux<false>(pc, s_pn[S(i, j + 1)], s_uy[S(i, j)]);
      const R f3 = donor_flux<false>(s_pn[S(i, j - 1)], pc, s_uy[S(i, j - 1)]);
      const R r = pc + ((-f0 + f1) + (-f2 + f3)) / 1;
      s_pk[S(i, j)] = r * R(1);
    }
    __syncthreads();
  }

  // ---- unlimited antidiffusive velocity: V_x on the faces owned by cells
  // (-2..T, -1..T), V_y on those owned by (-1..T, -2..T)
  {
    constexpr int W1 = kP2T + 3;
    for (int q = tid; q < W1 * W1; q += kP2Threads) {
      const int i = q / W1 - 2, j = q % W1 - 2;
      const bool st = a.storev && i >= 0 && i < kP2T && j >= 0 && j < kP2T;
      const int gx = x0 + i, gy = y0 + j;
      const bool in_v = st && gx >= a.vlo[0] && gx <= a.vhi[0] && gy >= a.vlo[1] && gy <= a.vhi[1];
      const long gp = L.org + (long)gx * L.st[0] + (long)gy * L.st[1];
      if (j >= -1) {
        // rotation (x, y): P(da, db) = psi at (i + da, j + db)
        auto P = [&](int da, int db, int) -> R { return s_pk[S(i + da, j + db)]; };
        const R s1 = s_uy[S(i + 1, j)] + s_uy[S(i, j)] + s_uy[S(i + 1, j - 1)] + s_uy[S(i, j - 1)];
        const R v = antidiff_core<R, G, 2, IGA, false, false>(s_ux[S(i, j)], s1, R(0), P);
        s_vx[S(i, j)] = v;
        if (in_v) a.V0[gp] = v;
      }
      if (i >= -1) {
        // rotation (y, x): P(da, db) = psi at (i + db, j + da)
        auto P = [&](int da, int db, int) -> R { return s_pk[S(i + db, j + da)]; };
        const R s1 = s_ux[S(i, j + 1)] + s_ux[S(i, j)] + s_ux[S(i - 1, j + 1)] + s_ux[S(i - 1, j)];
        const R v = antidiff_core<R, G, 2, IGA, false, false>(s_uy[S(i, j)], s1, R(0), P);
        s_vy[S(i, j)] = v;
        if (in_v) a.V1[gp] = v;
      }
    }
  }
  __syncthreads();

  // ---- beta_up / beta_dn over the tile plus 1 (FCT)
  if (FCT) {
    constexpr int WB = kP2T + 2;
    for (int q = tid; q < WB * WB; q += kP2Threads) {
      const int i = q / WB - 1, j = q % WB - 1;
      const R pc = s_pk[S(i, j)];
      const R xp = s_pk[S(i + 1, j)], xm = s_pk[S(i - 1, j)], yp = s_pk[S(i, j + 1)], ym = s_pk[S(i, j - 1)];
      R pmin, pmax;
      bounds_core<R, 2>(pmin, pmax, s_pn[S(i, j)], s_pn[S(i + 1, j)], s_pn[S(i - 1, j)], s_pn[S(i, j + 1)],
                        s_pn[S(i, j - 1)], R(0), R(0));
      R fxp, fxm, fyp, fym;
      if (IGA) {
        fxp = s_vx[S(i, j)]; fxm = s_vx[S(i - 1, j)]; fyp = s_vy[S(i, j)]; fym = s_vy[S(i, j - 1)];
      } else {
        fxp = donor_flux<false>(pc, xp, s_vx[S(i, j)]);
        fxm = donor_flux<false>(xm, pc, s_vx[S(i - 1, j)]);
        fyp = donor_flux<false>(pc, yp, s_vy[S(i, j)]);
        fym = donor_flux<false>(ym, pc, s_vy[S(i, j - 1)]);
      }
      R bu, bd;
      beta_core<R, G, 2, false>(bu, bd, pc, xp, xm, yp, ym, R(0), R(0), pmin, pmax, fxp, fxm, fyp, fym, R(0),
                                R(0));
      s_bu[S(i, j)] = bu;
      s_bd[S(i, j)] = bd;
    }
    __syncthreads();
  }

  // ---- limited fluxes and the donor-cell update of this thread's cell
  const int i = tid / kP2T, j = tid % kP2T;
  const int gx = x0 + i, gy = y0 + j;
  if (gx > a.hi[0] || gy > a.hi[1]) return;
  const R pc = s_pk[S(i, j)];
  const R xp = s_pk[S(i + 1, j)], xm = s_pk[S(i - 1, j)], yp = s_pk[S(i, j + 1)], ym = s_pk[S(i, j - 1)];
  R f0, f1, f2, f3;
  if (FCT) {
    const R bu0 = s_bu[S(i, j)], bd0 = s_bd[S(i, j)];
    const R mxp = mono_core<R, false, IGA>(s_vx[S(i, j)], bu0, bd0, s_bu[S(i + 1, j)], s_bd[S(i + 1, j)], pc, xp);
    const R mxm = mono_core<R, false, IGA>(s_vx[S(i - 1, j)], s_bu[S(i - 1, j)], s_bd[S(i - 1, j)], bu0, bd0, xm,
                                           pc);
    const R myp = mono_core<R, false, IGA>(s_vy[S(i, j)], bu0, bd0, s_bu[S(i, j + 1)], s_bd[S(i, j + 1)], pc, yp);
    const R mym = mono_core<R, false, IGA>(s_vy[S(i, j - 1)], s_bu[S(i, j - 1)], s_bd[S(i, j - 1)], bu0, bd0, ym,
                                           pc);
    f0 = IGA ? mxp : donor_flux<false>(pc, xp, mxp);
    f1 = IGA ? mxm : donor_flux<false>(xm, pc, mxm);
    f2 = IGA ? myp : donor_flux<false>(pc, yp, myp);
    f3 = IGA ? mym : donor_flux<false>(ym, pc, mym);
  } else {
    // the donor-cell pass with the velocity as Courant number (iga: as flux)
    const R c = pc * R(1);
    f0 = IGA ? s_vx[S(i, j)] : donor_flux<false>(c, xp * R(1), s_vx[S(i, j)]);
    f1 = IGA ? s_vx[S(i - 1, j)] : donor_flux<false>(xm * R(1), c, s_vx[S(i - 1, j)]);
    f2 = IGA ? s_vy[S(i, j)] : donor_flux<false>(c, yp * R(1), s_vy[S(i, j)]);
    f3 = IGA ? s_vy[S(i, j - 1)] : donor_flux<false>(ym * R(1), c, s_vy[S(i, j - 1)]);
  }
  const R r = pc + ((-f0 + f1) + (-f2 + f3)) / 1;
  a.out[L.org + (long)gx * L.st[0] + (long)gy * L.st[1]] = r * a.out_scale;
#undef s_pn
#undef s_pk
#undef s_ux
#undef s_uy
#undef s_vx
#undef s_vy
#undef s_bu
#undef s_bd
}

template <class R, class G, bool FCT, bool IGA>
__global__ void __launch_bounds__(kP2Threads) k_fz_pass2d(Pass2dArgs<R> a)
{
  fz_pass2d_block<R, G, FCT, IGA>(a, blockIdx.x);
}

// the 2D pass kernel covers standard or iga antidiff without abs/npa/pfc/khn
inline bool pass2d_supported(const mpdx_params &p, const Layout &L)
{
  if (L.nd != 2) return false;
  if (p.opts & (O_ABS | O_NPA | O_PFC | O_KHN)) return false;
  return L.n[0] >= 12 && L.n[1] >= 12;
}

}  // namespace mpdx

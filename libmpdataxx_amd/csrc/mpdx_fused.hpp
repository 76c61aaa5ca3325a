// Fused MPDATA kernels.
//
// One MPDATA pass costs three kernels instead of the per-pass path's ~15:
//   k_fz_donor      donor-cell pass: fluxes and the update in one sweep
//                   (donorcell_formulae.hpp:50-79,105-165); with iga the
//                   corrective velocity itself is the flux (mpdata_osc_3d.hpp:164-168)
//   k_fz_antidiff   all advector components of the antidiffusive velocity
//                   over their reference ranges (formulae_mpdata_3d.hpp:253-295)
//   k_fz_fct        FCT: psi_min/psi_max from psi^n, beta denominators from the
//                   donor fluxes of the unlimited velocity, beta_up/beta_dn,
//                   the limited velocity and the donor-cell update, all in
//                   registers/LDS (mpdata_fct_3d.hpp:31-110,
//                   formulae_mpdata_fct_3d.hpp:21-247)
// Between k_fz_antidiff and the next kernel the reference's own advector halo
// fills run on the boundary planes (same kernels as the per-pass path), so the
// fused path reproduces the reference for every input.  Scalar halos never
// exist in memory: psi is read through the per-dimension index map of the
// scalar halo fill (cyclic: period N-1; open: clamp to the edge).
#pragma once

#include <type_traits>

#include "mpdx_perpass.hpp"
#include "mpdx_solver.hpp"

struct mpdx_fused_state {
  void *psi_tmp[MPDX_MAX_EQNS] = {};
};

namespace mpdx {

// per-dimension index map of the scalar halo fill (bcond/cyclic_3d.hpp:29-33,
// bcond/open_3d.hpp:32-42): cyclic copies with period N-1, open copies the edge
struct BcMap {
  int n[3];
  unsigned char cyc[3][2];
};

__device__ inline int map_idx(const BcMap &m, int d, int i)
{
  const int N = m.n[d];
  if (i < 0) return m.cyc[d][0] ? i + (N - 1) : 0;
  if (i > N - 1) return m.cyc[d][1] ? i - (N - 1) : N - 1;
  return i;
}

// Scalar halo of width hw, materialized in one launch: every halo cell takes
// the value of the interior cell that the reference's xchng_sclr
// (solver_3d.hpp:47-67, solver_2d.hpp:40-56) leaves there.  Its fills run
// dimension by dimension over ranges widened by the halo, so a halo cell
// ends up with the composition of the per-dimension rules: cyclic wrap by
// N-1 (bcond/cyclic_3d.hpp:29-33) or open edge copy (bcond/open_3d.hpp:
// 32-42).  Sources are interior cells, so the launch has no ordering hazard.
// One launch covers every slab (a box list): x slabs over the full extended
// (y, z) range, y slabs over interior x, z slabs over interior (x, y) — or,
// with z-boundary tiles, only the bands of (x, y) rows that the boundary
// shell reads.
template <class R>
__device__ __forceinline__ void fz_halo_cell(R *a, const Layout &L, const BcMap &m, int i, int j, int k)
{
  const int si = map_idx(m, 0, i), sj = map_idx(m, 1, j), sk = L.nd == 3 ? map_idx(m, 2, k) : 0;
  a[L.org + i * L.st[0] + j * L.st[1] + k * L.st[2]] = a[L.org + si * L.st[0] + sj * L.st[1] + sk * L.st[2]];
}

template <class R>
__global__ void __launch_bounds__(256) k_fz_halo(R *a, Layout L, BcMap m, BoxList bl)
{
  int q, i, j, k;
  if (!boxlist_idx(bl, q, i, j, k)) return;
  fz_halo_cell(a, L, m, i, j, k);
}

// element index of cell (i, j, k); the (j, k) part in 32-bit arithmetic
// (a plane of the halo-extended array has fewer than 2^31 elements)
__device__ __forceinline__ long cell_index(const Layout &L, int i, int j, int k)
{
  return L.org + (long)i * L.st[0] + (long)(j * (int)L.st[1] + k * (int)L.st[2]);
}

// A at the cell (dx, dy, dz) away from element p (constant offsets fold into
// stride multiples); cells up to hw outside the domain come from the
// materialized halo
template <class R>
__device__ __forceinline__ R rel(const R *A, long p, const Layout &L, int dx, int dy, int dz)
{
  return A[p + dx * L.st[0] + dy * L.st[1] + dz * L.st[2]];
}

// ---------------------------------------------------------------- donor-cell pass

// psi + the flux divergence (donorcell_sum, donorcell_formulae.hpp:105-165):
// f[0..5] are the fluxes through the +x, -x, +y, -y, +z, -z faces.  Without
// khn the reference's parenthesisation; with khn (opts::khn) the
// compensated sum of psi and the signed fluxes in the reference's order,
// each step kahan_add (formulae/kahan_sum.hpp:13-27); the switch is uniform
// across a launch
template <class R, int ND>
__device__ __forceinline__ R donor_sum(R pc, const R *f, int khn)
{
  if (!khn) {
    if (ND == 3) return pc + ((-f[0] + f[1]) + (-f[2] + f[3]) + (-f[4] + f[5])) / 1;
    return pc + ((-f[0] + f[1]) + (-f[2] + f[3])) / 1;
  }
  R sum = 0, c = 0;
  auto add = [&](R in) {
    const R y = in - c;
    const R t = sum + y;
    c = (t - sum) - y;
    sum = t;
  };
  add(pc);
  for (int q = 0; q < 2 * ND; ++q) add(q % 2 == 0 ? -f[q] / 1 : f[q] / 1);
  return sum;
}

template <class R, bool NPA>
__device__ __forceinline__ void fz_donor_cell(R *out, const R *in, const R *C0, const R *C1, const R *C2,
                                              const Layout &L, R in_scale, R out_scale, int flux_is_C, int khn,
                                              int i, int j, int k)
{
  const long p = cell_index(L, i, j, k);
  const R pc = in[p] * in_scale;
  R f[6];
  if (flux_is_C) {
    f[0] = C0[p]; f[1] = rel(C0, p, L, -1, 0, 0);
    f[2] = C1[p]; f[3] = rel(C1, p, L, 0, -1, 0);
    if (L.nd == 3) { f[4] = C2[p]; f[5] = rel(C2, p, L, 0, 0, -1); }
  } else {
    const R xp = rel(in, p, L, 1, 0, 0) * in_scale, xm = rel(in, p, L, -1, 0, 0) * in_scale;
    const R yp = rel(in, p, L, 0, 1, 0) * in_scale, ym = rel(in, p, L, 0, -1, 0) * in_scale;
    f[0] = donor_flux<NPA>(pc, xp, C0[p]);
    f[1] = donor_flux<NPA>(xm, pc, rel(C0, p, L, -1, 0, 0));
    f[2] = donor_flux<NPA>(pc, yp, C1[p]);
    f[3] = donor_flux<NPA>(ym, pc, rel(C1, p, L, 0, -1, 0));
    if (L.nd == 3) {
      const R zp = rel(in, p, L, 0, 0, 1) * in_scale, zm = rel(in, p, L, 0, 0, -1) * in_scale;
      f[4] = donor_flux<NPA>(pc, zp, C2[p]);
      f[5] = donor_flux<NPA>(zm, pc, rel(C2, p, L, 0, 0, -1));
    }
  }
  const R r = L.nd == 3 ? donor_sum<R, 3>(pc, f, khn) : donor_sum<R, 2>(pc, f, khn);
  out[p] = r * out_scale;
}

template <class R, bool NPA>
__global__ void __launch_bounds__(256)
k_fz_donor(R *out, const R *in, const R *C0, const R *C1, const R *C2, Layout L,
           R in_scale, R out_scale, int flux_is_C, int khn, Box bx)
{
  // cells of the box bx (last coordinate fastest)
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  fz_donor_cell<R, NPA>(out, in, C0, C1, C2, L, in_scale, out_scale, flux_is_C, khn, i, j, k);
}

// ---------------------------------------------------------------- antidiffusive velocity

// Antidiffusive velocity at one face from its psi stencil and the two
// tangential advector sums (formulae_mpdata_3d.hpp:253-295,
// formulae_mpdata_2d.hpp:241-278, psi_3d.hpp:261-464).  P(da, db, dc) returns
// psi at the rotated offset (da along the face normal, db / dc along the
// first / second tangential direction of idxperm::pi<d>); s1, s2 are the
// four-point sums whose quarter is the reference's GC1_bar / GC2_bar.
// The reference's factors 1/2 (|u|/2, u/2), 1/4 (the averages) and 2 (ndx)
// are powers of two, so they are applied once at the end: |u|(1-|u|) frac
// - u (s1 ndy + s2 ndz) / 8 (2D: - u s1 ndy / 8), with |u|(1-|u|) as one
// fused multiply-add.
// PAIRS (the interior kernel): the four-point sums of the tangential ratios
// as sums of their two pairs, (P+ + P-) +- (M+ + M-), four additions where
// the reference's left-to-right order takes five (rounding differs in the
// last bit)
template <class R, class G, int ND, bool IGA, bool ABS, bool PFC, class PF, bool PAIRS = false>
__device__ inline R antidiff_core(R u, R s1, R s2, PF P)
{
  auto Q = [&](int da, int db, int dc) -> R {
    const R v = P(da, db, dc);
    return ABS ? fabs(v) : v;
  };
  const R au = fabs(u);
  if (IGA) {
    const R ndx = ND == 3 ? P(1, 0, 0) - P(0, 0, 0) : 2 * (P(1, 0, 0) - P(0, 0, 0)) / (1 + 1);
    const R ndy = (P(1, 1, 0) + P(0, 1, 0) - P(1, -1, 0) - P(0, -1, 0)) / 4;
    const R g1 = s1 / 4;
    if (ND == 3) {
      const R g2 = s2 / 4;
      const R ndz = (P(1, 0, 1) + P(0, 0, 1) - P(1, 0, -1) - P(0, 0, -1)) / 4;
      return au / 2 * (1 - au / 1) * ndx - u / 2 * (g1 * ndy + g2 * ndz) / 1;
    }
    return au / 2 * (1 - au / 1) * ndx - u * g1 / (2 * 1) * ndy;
  }
  const R a1 = Q(1, 0, 0), a0 = Q(0, 0, 0);
  const R fx = frac<PFC, G>(a1 - a0, a1 + a0);
  const R b11 = Q(1, 1, 0), b01 = Q(0, 1, 0), b1m = Q(1, -1, 0), b0m = Q(0, -1, 0);
  auto ratio = [](R pp, R p0, R mp, R m0) -> R {
    if constexpr (PAIRS) {
      const R P = pp + p0, M = mp + m0;
      return frac<PFC, G>(P - M, P + M);
    } else {
      return frac<PFC, G>(pp + p0 - mp - m0, pp + p0 + mp + m0);
    }
  };
  const R ndy = ratio(b11, b01, b1m, b0m);
  const R w = fma(-au, au, au) * fx;
  if (ND == 3) {
    const R c11 = Q(1, 0, 1), c01 = Q(0, 0, 1), c1m = Q(1, 0, -1), c0m = Q(0, 0, -1);
    const R ndz = ratio(c11, c01, c1m, c0m);
    return fma(-(u * R(0.125)), fma(s2, ndz, s1 * ndy), w);
  }
  return fma(-(u * R(0.125)), s1 * ndy, w);
}

// V_D at face (a+1/2 along D) of the cell at element p: the rotation
// idxperm::pi<D> (formulae/idxperm.hpp:96-125) maps (a, b, c) to
// (x, y, z), (y, z, x), (z, x, y) in 3D and (x, y), (y, x) in 2D; Ud, U1, U2
// are GC_unco[D], [D+1], [D+2]
template <class R, class G, int ND, int D, bool IGA, bool ABS, bool PFC>
__device__ inline R antidiff_at(const R *psi, const R *Ud, const R *U1, const R *U2, const Layout &L,
                                long p)
{
  // cell offset of the rotated (a, b, c) offset
  auto off = [&](const R *A, int da, int db, int dc) -> R {
    if (ND == 2) return D == 0 ? rel(A, p, L, da, db, 0) : rel(A, p, L, db, da, 0);
    if (D == 0) return rel(A, p, L, da, db, dc);
    if (D == 1) return rel(A, p, L, dc, da, db);
    return rel(A, p, L, db, dc, da);
  };
  auto P = [&](int da, int db, int dc) -> R { return off(psi, da, db, dc); };
  const R s1 = off(U1, 1, 0, 0) + U1[p] + off(U1, 1, -1, 0) + off(U1, 0, -1, 0);
  if (ND == 3) {
    const R s2 = off(U2, 1, 0, 0) + U2[p] + off(U2, 1, 0, -1) + off(U2, 0, 0, -1);
    return antidiff_core<R, G, 3, IGA, ABS, PFC>(Ud[p], s1, s2, P);
  }
  return antidiff_core<R, G, 2, IGA, ABS, PFC>(Ud[p], s1, R(0), P);
}

// all components over the box [-1, n-1]^nd; component d is written where the
// other coordinates are interior (the reference's im x j x k ranges)
template <class R, class G, int ND, bool IGA, bool ABS, bool PFC>
__device__ __forceinline__ void fz_antidiff_cell(R *V0, R *V1, R *V2, const R *psi, const R *U0, const R *U1,
                                                 const R *U2, const Layout &L, int i, int j, int k)
{
  const bool in0 = i >= 0, in1 = j >= 0, in2 = ND == 2 || k >= 0;
  const long p = cell_index(L, i, j, k);
  if (ND == 3) {
    if (in1 && in2) V0[p] = antidiff_at<R, G, 3, 0, IGA, ABS, PFC>(psi, U0, U1, U2, L, p);
    if (in2 && in0) V1[p] = antidiff_at<R, G, 3, 1, IGA, ABS, PFC>(psi, U1, U2, U0, L, p);
    if (in0 && in1) V2[p] = antidiff_at<R, G, 3, 2, IGA, ABS, PFC>(psi, U2, U0, U1, L, p);
  } else {
    if (in1) V0[p] = antidiff_at<R, G, 2, 0, IGA, ABS, PFC>(psi, U0, U1, U1, L, p);
    if (in0) V1[p] = antidiff_at<R, G, 2, 1, IGA, ABS, PFC>(psi, U1, U0, U0, L, p);
  }
}

template <class R, class G, int ND, bool IGA, bool ABS, bool PFC>
__global__ void __launch_bounds__(256)
k_fz_antidiff(R *V0, R *V1, R *V2, const R *psi, const R *U0, const R *U1, const R *U2,
              Layout L, Box bx)
{
  // bx: sub-box of the owner cells [-1, n-1]^nd (last coordinate fastest)
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  fz_antidiff_cell<R, G, ND, IGA, ABS, PFC>(V0, V1, V2, psi, U0, U1, U2, L, i, j, k);
}

// ---------------------------------------------------------------- FCT + donor-cell

// One thread per interior cell; the block owns a brick of cells and stages
// beta over the brick plus a one-cell frame in LDS, so each beta is computed
// once per block.  3D brick 2 x 4 x 32 (x, y, z); 2D brick 8 x 32 (x, y).
// SHAPE picks the brick for a slab that is thin along x (0), y (1) or z (2)
template <int ND, int SHAPE> struct FzBrick;
template <> struct FzBrick<3, 0> { static constexpr int TX = 2, TY = 4, TZ = 32; };
template <> struct FzBrick<3, 1> { static constexpr int TX = 8, TY = 2, TZ = 16; };
template <> struct FzBrick<3, 2> { static constexpr int TX = 8, TY = 16, TZ = 2; };
template <> struct FzBrick<2, 0> { static constexpr int TX = 8, TY = 32, TZ = 1; };
template <> struct FzBrick<2, 1> { static constexpr int TX = 128, TY = 2, TZ = 1; };
template <> struct FzBrick<2, 2> { static constexpr int TX = 8, TY = 32, TZ = 1; };

template <int ND, int SHAPE> constexpr int fz_brick_threads()
{
  return FzBrick<ND, SHAPE>::TX * FzBrick<ND, SHAPE>::TY * FzBrick<ND, SHAPE>::TZ;
}

// beta_up / beta_dn of one cell from psi^k (centre + 6 neighbours), the
// psi^n bounds and the six face fluxes of the unlimited velocity
// (formulae_mpdata_fct_3d.hpp:21-145, formulae_mpdata_fct_2d.hpp:21-120)
template <class R, class G, int ND, bool NPA>
__device__ inline void beta_core(R &bup, R &bdn, R pc, R xp, R xm, R yp, R ym, R zp, R zm, R pmin,
                                 R pmax, R fxp, R fxm, R fyp, R fym, R fzp, R fzm)
{
  R up_nom, dn_nom, up_den, dn_den;
  if (ND == 3) {
    up_nom = (smax(pmax, smax(pc, smax(xp, smax(xm, smax(yp, smax(ym, smax(zp, zm))))))) - pc) * 1;
    dn_nom = (pc - smin(pmin, smin(yp, smin(xm, smin(pc, smin(xp, smin(zp, smin(zm, ym)))))))) * 1;
    up_den = (pospart<NPA>(fxm) - negpart<NPA>(fxp)) + (pospart<NPA>(fym) - negpart<NPA>(fyp))
           + (pospart<NPA>(fzm) - negpart<NPA>(fzp));
    dn_den = (pospart<NPA>(fxp) - negpart<NPA>(fxm)) + (pospart<NPA>(fyp) - negpart<NPA>(fym))
           + (pospart<NPA>(fzp) - negpart<NPA>(fzm));
  } else {
    up_nom = (smax(pmax, smax(yp, smax(xm, smax(pc, smax(xp, ym))))) - pc) * 1;
    dn_nom = (pc - smin(pmin, smin(yp, smin(xm, smin(pc, smin(xp, ym)))))) * 1;
    up_den = (pospart<NPA>(fxm) - negpart<NPA>(fxp)) + (pospart<NPA>(fym) - negpart<NPA>(fyp));
    dn_den = (pospart<NPA>(fxp) - negpart<NPA>(fxm)) + (pospart<NPA>(fyp) - negpart<NPA>(fym));
  }
  bup = fct_frac<G>(up_nom, up_den);
  bdn = fct_frac<G>(dn_nom, dn_den);
}

// psi_min / psi_max of psi^n over the 7-point (5-point) stencil
// (mpdata_fct_3d.hpp:31-55, mpdata_fct_2d.hpp:31-46)
template <class R, int ND>
__device__ inline void bounds_core(R &pmin, R &pmax, R c, R a, R b, R d, R e, R f, R g)
{
  // a = x+1, b = x-1, d = y+1, e = y-1, f = z+1, g = z-1
  if (ND == 3) {
    pmin = smin(smin(smin(smin(smin(smin(c, a), b), d), e), f), g);
    pmax = smax(smax(smax(smax(smax(smax(c, a), b), d), e), f), g);
  } else {
    pmin = smin(smin(smin(smin(d, b), c), a), e);
    pmax = smax(smax(smax(smax(d, b), c), a), e);
  }
}

// limited velocity (GC_mono) on a face between cells l and r
// (formulae_mpdata_fct_3d.hpp:147-247)
template <class R, bool ABS, bool IGA>
__device__ inline R mono_core(R v, R bul, R bdl, R bur, R bdr, R psl, R psr)
{
  const R m_dn_up = smin(R(1), smin(bdl, bur));
  const R m_up_dn = smin(R(1), smin(bul, bdr));
  R lim;
  if (!ABS || IGA) lim = v > 0 ? m_dn_up : m_up_dn;
  else if (v > 0) lim = psl > 0 ? m_dn_up : m_up_dn;
  else lim = psr > 0 ? m_up_dn : m_dn_up;
  return v * lim;
}

template <class R, class G, bool NPA, bool ABS>
__device__ inline void beta_at(R &bup, R &bdn, const R *psi, const R *psin, const R *V0, const R *V1,
                               const R *V2, const Layout &L, int i, int j, int k, R in_scale, bool iga)
{
  const int nd = L.nd;
  const long p = cell_index(L, i, j, k);
  auto P = [&](int a, int b, int c) { return rel(psi, p, L, a, b, c); };
  auto Q = [&](int a, int b, int c) { return rel(psin, p, L, a, b, c) * in_scale; };
  const R pc = psi[p];
  const R xp = P(1, 0, 0), xm = P(-1, 0, 0), yp = P(0, 1, 0), ym = P(0, -1, 0);
  R zp = 0, zm = 0;
  if (nd == 3) { zp = P(0, 0, 1); zm = P(0, 0, -1); }
  R pmin, pmax;
  if (nd == 3)
    bounds_core<R, 3>(pmin, pmax, Q(0, 0, 0), Q(1, 0, 0), Q(-1, 0, 0), Q(0, 1, 0),
                      Q(0, -1, 0), Q(0, 0, 1), Q(0, 0, -1));
  else
    bounds_core<R, 2>(pmin, pmax, Q(0, 0, 0), Q(1, 0, 0), Q(-1, 0, 0), Q(0, 1, 0),
                      Q(0, -1, 0), R(0), R(0));
  R fxp, fxm, fyp, fym, fzp = 0, fzm = 0;
  if (iga) {
    fxp = V0[p]; fxm = rel(V0, p, L, -1, 0, 0); fyp = V1[p]; fym = rel(V1, p, L, 0, -1, 0);
    if (nd == 3) { fzp = V2[p]; fzm = rel(V2, p, L, 0, 0, -1); }
  } else {
    fxp = donor_flux<NPA>(pc, xp, V0[p]);
    fxm = donor_flux<NPA>(xm, pc, rel(V0, p, L, -1, 0, 0));
    fyp = donor_flux<NPA>(pc, yp, V1[p]);
    fym = donor_flux<NPA>(ym, pc, rel(V1, p, L, 0, -1, 0));
    if (nd == 3) {
      fzp = donor_flux<NPA>(pc, zp, V2[p]);
      fzm = donor_flux<NPA>(zm, pc, rel(V2, p, L, 0, 0, -1));
    }
  }
  if (nd == 3) beta_core<R, G, 3, NPA>(bup, bdn, pc, xp, xm, yp, ym, zp, zm, pmin, pmax, fxp, fxm, fyp, fym, fzp, fzm);
  else beta_core<R, G, 2, NPA>(bup, bdn, pc, xp, xm, yp, ym, zp, zm, pmin, pmax, fxp, fxm, fyp, fym, fzp, fzm);
}

// brick (gz, gy, gx) of the box: gz along z (3D), gy along y, gx along x
template <class R, class G, int ND, int SHAPE, bool NPA, bool ABS, bool IGA>
__device__ __forceinline__ void fz_fct_block(R *out, const R *psi, const R *psin, const R *V0, const R *V1,
                                             const R *V2, const Layout &L, R in_scale, R out_scale, int khn,
                                             const Box &bx, unsigned gz, unsigned gy, unsigned gx)
{
  constexpr int TX = FzBrick<ND, SHAPE>::TX, TY = FzBrick<ND, SHAPE>::TY, TZ = FzBrick<ND, SHAPE>::TZ;
  constexpr int BX = TX + 2, BY = TY + 2, BZ = ND == 3 ? TZ + 2 : 1;
  __shared__ R sbu[BX * BY * BZ], sbd[BX * BY * BZ];
  const int i0 = bx.lo[0] + gx * TX, j0 = bx.lo[1] + gy * TY;
  const int k0 = ND == 3 ? bx.lo[2] + gz * TZ : 0;
  const int tid = threadIdx.x;
  // beta over the brick plus a one-cell frame (cells -1..n of the domain)
  for (int q = tid; q < BX * BY * BZ; q += blockDim.x) {
    const int bz = q % BZ, by = (q / BZ) % BY, bx = q / (BZ * BY);
    const int i = i0 - 1 + bx, j = j0 - 1 + by, k = ND == 3 ? k0 - 1 + bz : 0;
    R bu = 0, bd = 0;
    const bool inb = i >= -1 && i <= L.n[0] && j >= -1 && j <= L.n[1] &&
                     (ND == 2 || (k >= -1 && k <= L.n[2]));
    if (inb) beta_at<R, G, NPA, ABS>(bu, bd, psi, psin, V0, V1, V2, L, i, j, k, in_scale, IGA);
    sbu[q] = bu;
    sbd[q] = bd;
  }
  __syncthreads();
  const int tz = tid % TZ, ty = (tid / TZ) % TY, tx = tid / (TZ * TY);
  const int i = i0 + tx, j = j0 + ty, k = ND == 3 ? k0 + tz : 0;
  if (i > bx.lo[0] + bx.ext[0] - 1 || j > bx.lo[1] + bx.ext[1] - 1 ||
      (ND == 3 && k > bx.lo[2] + bx.ext[2] - 1)) return;
  auto B = [&](const R *sb, int di, int dj, int dk) -> R {
    return sb[((tx + 1 + di) * BY + (ty + 1 + dj)) * BZ + (ND == 3 ? tz + 1 + dk : 0)];
  };
  const long p = cell_index(L, i, j, k);
  const R pc = psi[p];
  const R bu0 = B(sbu, 0, 0, 0), bd0 = B(sbd, 0, 0, 0);
  auto mono = [&](R v, R bul, R bdl, R bur, R bdr, R psl, R psr) -> R {
    return mono_core<R, ABS, IGA>(v, bul, bdl, bur, bdr, psl, psr);
  };
  R f[6];
  {
    const R xp = rel(psi, p, L, 1, 0, 0), xm = rel(psi, p, L, -1, 0, 0);
    const R mp = mono(V0[p], bu0, bd0, B(sbu, 1, 0, 0), B(sbd, 1, 0, 0), pc, xp);
    const R mm = mono(rel(V0, p, L, -1, 0, 0), B(sbu, -1, 0, 0), B(sbd, -1, 0, 0), bu0, bd0, xm, pc);
    f[0] = IGA ? mp : donor_flux<NPA>(pc, xp, mp);
    f[1] = IGA ? mm : donor_flux<NPA>(xm, pc, mm);
  }
  {
    const R yp = rel(psi, p, L, 0, 1, 0), ym = rel(psi, p, L, 0, -1, 0);
    const R mp = mono(V1[p], bu0, bd0, B(sbu, 0, 1, 0), B(sbd, 0, 1, 0), pc, yp);
    const R mm = mono(rel(V1, p, L, 0, -1, 0), B(sbu, 0, -1, 0), B(sbd, 0, -1, 0), bu0, bd0, ym, pc);
    f[2] = IGA ? mp : donor_flux<NPA>(pc, yp, mp);
    f[3] = IGA ? mm : donor_flux<NPA>(ym, pc, mm);
  }
  R r;
  if (ND == 3) {
    const R zp = rel(psi, p, L, 0, 0, 1), zm = rel(psi, p, L, 0, 0, -1);
    const R mp = mono(V2[p], bu0, bd0, B(sbu, 0, 0, 1), B(sbd, 0, 0, 1), pc, zp);
    const R mm = mono(rel(V2, p, L, 0, 0, -1), B(sbu, 0, 0, -1), B(sbd, 0, 0, -1), bu0, bd0, zm, pc);
    f[4] = IGA ? mp : donor_flux<NPA>(pc, zp, mp);
    f[5] = IGA ? mm : donor_flux<NPA>(zm, pc, mm);
    r = donor_sum<R, 3>(pc, f, khn);
  } else {
    r = donor_sum<R, 2>(pc, f, khn);
  }
  out[p] = r * out_scale;
}

template <class R, class G, int ND, int SHAPE, bool NPA, bool ABS, bool IGA>
__global__ void __launch_bounds__((fz_brick_threads<ND, SHAPE>()))
k_fz_fct(R *out, const R *psi, const R *psin, const R *V0, const R *V1, const R *V2, Layout L,
         R in_scale, R out_scale, int khn, Box bx)
{
  fz_fct_block<R, G, ND, SHAPE, NPA, ABS, IGA>(out, psi, psin, V0, V1, V2, L, in_scale, out_scale, khn, bx,
                                               blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------- interior corrective pass (3D)

// One whole corrective pass (antidiffusive velocity, FCT limiter, donor-cell
// update) for cells at distance >= 2 from every boundary, where no halo rule
// touches any value the pass reads, so no advector halo fill is needed.  With
// open bounds on both z sides (InteriorArgs::zb) the z tiles also cover the
// cells next to the z boundaries and apply the open z halo rules in the tile
// (halo columns, below).  A block owns a (Y-4) x (Z-4) column of output cells and marches along x
// through a chunk of output planes; each thread owns one (y, z) column of the
// (Y x Z) tile, which includes a 2-cell frame.  At step q (newest psi^k plane
// p = x0 - 2 + q) it computes the velocity on plane p-1, beta on plane p-1,
// the limited fluxes and writes psi^{k+1} on plane p-2.
//
// With FUSE0 (the first corrective pass) the kernel reads psi^n and GC
// instead of psi^1 and computes psi^1 itself with the donor-cell formula
// (donorcell_formulae.hpp:105-165): psi^n and GC run one plane ahead, and
// step q produces psi^1 on plane p+1 after its own output.  The donor needs
// psi^n and GC_y / GC_z one cell outside the tile, which a few threads load
// into the LDS frame.
//
// psi and psi^n planes, the in-plane advector and velocity components, beta
// and the y/z fluxes go through LDS rings; a thread's own column stays in
// registers.  The march is unrolled by 6 (the period of the 3- and 2-plane
// rings) so every ring slot is a compile-time constant.  Each LDS plane is
// the tile plus a one-cell frame, so every neighbour read is the thread's
// own LDS address plus an immediate offset; frame cells that nobody loads
// hold zeros and reach only frame results, which are discarded.
// Tile of TYV (y) x 32 (z) threads.  FP32 tiles are 32 high (one
// 1024-thread block per CU, 72-99 KB of LDS) or 16 high (512 threads, two
// blocks per CU); FP64 tiles are 24 high (768 threads, 12 waves: one block
// per CU, 110-151 KB of LDS; 32-high rings would need up to 198 KB, more
// than a CU's 160 KB) or 16 high (below).
template <int TYV> struct ITile {
  static constexpr int Y = TYV, Z = 32;
  static constexpr int PW = Z + 2, PLANE = (Y + 2) * PW;
  static constexpr int NFRAME = 3 * Z + 3 * Y;   // FUSE0 frame loads per plane
};
// the default tile height of a kernel variant: FP32 32 rows; FP64 24, but
// 16 for the first pass that is also the last (FUSE0 without STOREV, n_iters
// = 2), which at 24 rows needs more than the 168 VGPRs three waves per SIMD
// allow and would spill
template <class R, bool FUSE0 = false, bool STOREV = true> constexpr int itile_y_default()
{
  return sizeof(R) == 8 ? (FUSE0 && !STOREV ? 16 : 24) : 32;
}

// LDS planes of the interior kernel: the psi^k and U_x rings (3 planes),
// the U_y, U_z and psi^n rings (2 planes, 3 with FUSE0), the limited y and z
// fluxes (FCT: one plane each, read by the update of the plane before they
// are overwritten; without FCT a 2-plane ring each) and, with FCT, V_y,
// V_z, beta-up and beta-down (one plane each)
template <bool FCT, bool FUSE0> struct ILds {
  static constexpr int RU = FUSE0 ? 3 : 2, RF = FCT ? 1 : 2;
  static constexpr int PSI = 0, UX = 3, UY = 6, UZ = UY + RU, PN = UZ + RU, FY = PN + RU, FZ = FY + RF,
                       VY = FZ + RF, VZ = VY + 1, BU = VY + 2, BD = VY + 3, N = VY + (FCT ? 4 : 0);
};

// LDS elements of an interior block: the planes, and with FUSE0 one more
// tile of psi^1 columns
template <bool FCT, bool FUSE0, int TYV> constexpr int interior_lds_elems()
{
  return ILds<FCT, FUSE0>::N * ITile<TYV>::PLANE + (FUSE0 ? ITile<TYV>::Y * ITile<TYV>::Z : 0);
}

template <class R>
struct InteriorArgs {
  R *out;
  const R *psi, *psin;  // psi^k (unused with FUSE0) and psi^n
  const R *U0, *U1, *U2;
  R *V0, *V1, *V2;      // unlimited velocity out (next pass's GC_unco), may be null
  Layout L;
  int lo[3], hi[3];     // output cells, inclusive; lo >= 2, hi <= n-3
  int vlo[3], vhi[3];   // cells whose velocity is stored (STOREV); the rest is the boundary shell's
  int nb[3];            // tiles along z, y and x chunks
  int chunk;            // output planes per x chunk
  R in_scale, out_scale;
  // zb != 0: the tiles also cover the z-boundary cells (lo[2] = 0,
  // hi[2] = n-1); the last z tile starts at zs_max so that it ends on the
  // last halo column.  zb = 1: open bounds on both z sides.  zb = 2: cyclic
  // bounds on both z sides with planes z = 0 and z = n-1 of psi and of the
  // tangential advector components equal (the duplicated edge of a period
  // of n-1 cells)
  int zb, zs_max;
  // with the kernels' WXY (and zb = 2): cyclic bounds in x and y too, with
  // the duplicated planes equal there as well; the tiles also cover every x
  // and y cell (lo = 0, hi = n-1; the last y tile starts at ys_max) and read
  // x and y halo planes at their periodic images, so no boundary shell
  // remains
  int ys_max;
};

// max / min without the std::max argument-order semantics (they differ only
// for NaNs and the sign of zero, which no result of this pass depends on);
// FP32 uses IEEE 754-2019 maximum / minimum, which gfx950 executes as one
// three-operand v_maximum3 / v_minimum3 and needs no NaN quieting of inputs
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ double fmx(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ double fmn(double a, double b) { return fmin(a, b); }

template <class R, class G>
__device__ __forceinline__ void ibeta(R &bup, R &bdn, R pc, R xp, R xm, R yp, R ym, R zp, R zm, R pmin,
                                      R pmax, R fxp, R fxm, R fyp, R fym, R fzp, R fzm)
{
  const R mx = fmx(fmx(fmx(pmax, pc), fmx(xp, xm)), fmx(fmx(yp, ym), fmx(zp, zm)));
  const R mn = fmn(fmn(fmn(pmin, pc), fmn(xp, xm)), fmn(fmn(yp, ym), fmn(zp, zm)));
  // sums of non-negative terms: no cancellation (rewriting them as (A +- B)/2
  // with A = sum |f| and B the net inflow cancels when every face flows the
  // same way, and a denominator that should be 0 comes out as +-1 ulp of A)
  const R z = 0;
  const R up_den = (fmx(z, fxm) - fmn(z, fxp)) + (fmx(z, fym) - fmn(z, fyp)) + (fmx(z, fzm) - fmn(z, fzp));
  const R dn_den = (fmx(z, fxp) - fmn(z, fxm)) + (fmx(z, fyp) - fmn(z, fym)) + (fmx(z, fzp) - fmn(z, fzm));
  bup = fct_frac<G>(mx - pc, up_den);
  bdn = fct_frac<G>(pc - mn, dn_den);
}

template <class R>
__device__ __forceinline__ R imono(R v, R bul, R bdl, R bur, R bdr)
{
  const R m_dn_up = fmn(R(1), fmn(bdl, bur));
  const R m_up_dn = fmn(R(1), fmn(bul, bdr));
  return v * (v > 0 ? m_dn_up : m_up_dn);
}

// donor-cell flux of the limited velocity m = v * limiter (limiter >= 0):
// upwind by the sign of v, which is the sign of m wherever m != 0 (the
// flux is then 0 either way, up to the sign of a zero)
template <class R> __device__ __forceinline__ R iflux_v(R l, R r, R m, R v)
{
  return m * (v > 0 ? l : r);
}

// donor-cell flux pospart(c) l + negpart(c) r: the same value (up to the
// sign of a zero) as c times the upwind psi
template <class R> __device__ __forceinline__ R iflux(R l, R r, R c)
{
  return c * (c > 0 ? l : r);
}

template <int N> using ic = std::integral_constant<int, N>;

// Loads and stores through a buffer resource built from a block-uniform
// base (scalar registers) with a 32-bit per-thread byte offset: no 64-bit
// vector address arithmetic per access, and the running bases advance with
// scalar adds.  Range 0xffffffff (no range check), dword 3 as for gfx9.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const char *g)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(g), (short)0, (int)0xffffffff, 0x00020000);
}
template <class R> __device__ __forceinline__ R buf_ld(const char *g, unsigned off)
{
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(g), (int)off, 0, 0));
  else
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(buf_rsrc(g), (int)off, 0, 0));
}
template <class R> __device__ __forceinline__ void buf_st(char *g, unsigned off, R v)
{
  if constexpr (sizeof(R) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), buf_rsrc(g), (int)off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(
                                              buf_rsrc(g), 0, 0, 0)), v),
                                          buf_rsrc(g), (int)off, 0, 0);
}

// Waves per SIMD an interior kernel is compiled for: FP32 tiles 24 rows
// high (768 threads) run two blocks per CU, 6 waves per SIMD, which caps
// them at 80 VGPRs; 1 leaves the others to their launch bounds.  Applied in
// the translation unit that instantiates the FP32 default-guard kernels
// (MPDX_INTERIOR_TU); the others carry no such attribute.
template <class R, int TYV> constexpr int interior_waves_per_eu()
{
  return sizeof(R) == 4 && TYV == 24 ? 6 : 1;
}
#ifdef MPDX_INTERIOR_TU
#define MPDX_INTERIOR_ATTR __attribute__((amdgpu_waves_per_eu(interior_waves_per_eu<R, TYV>())))
#else
#define MPDX_INTERIOR_ATTR
#endif
// block blk of the launch over a's tiles (a launch of 8 * ceil(tiles / 8)
// blocks; blk % 8 is the XCD the block runs on when the launch starts on a
// multiple of 8)
// the LDS array of an interior block, one per element type and size: a
// kernel whose blocks run one of several variants (InteriorMixTask) gives
// them all the largest size and so holds one array
template <class R, int N> __shared__ R interior_lds[N];

// LDS_N: the size of the block's LDS array (0: interior_lds_elems of this
// variant)
template <class R, class G, bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV, bool WXY, int LDS_N = 0>
__device__ __forceinline__ void fz_interior_block(InteriorArgs<R> a, unsigned blk)
{
  constexpr int TY = ITile<TYV>::Y, TZ = ITile<TYV>::Z, PW = ITile<TYV>::PW, PL = ITile<TYV>::PLANE;
  using IL = ILds<FCT, FUSE0>;
  constexpr int IL_PSI = IL::PSI, IL_UX = IL::UX, IL_UY = IL::UY, IL_UZ = IL::UZ, IL_PN = IL::PN,
                IL_FY = IL::FY, IL_FZ = IL::FZ, IL_VY = IL::VY, IL_VZ = IL::VZ, IL_BU = IL::BU, IL_BD = IL::BD,
                IL_N = IL::N;
  static_assert(LDS_N == 0 || LDS_N >= interior_lds_elems<FCT, FUSE0, TYV>(), "LDS array too small");
  constexpr int LN = LDS_N ? LDS_N : interior_lds_elems<FCT, FUSE0, TYV>();

  // XCD-aware tile order: workgroup w runs on XCD w % 8, and XCD x takes the
  // x-th contiguous run of tiles (z fastest), so neighbouring tiles that
  // share halo cache lines run on the same XCD at the same time
  const int nbz = a.nb[0], nby = a.nb[1];
  const int ntile = nbz * nby * a.nb[2];
  const int per = (ntile + 7) / 8;
  const int t = (int)(blk & 7) * per + (int)(blk >> 3);
  if (t >= ntile) return;
  const int bz = t % nbz, by = (t / nbz) % nby, bxc = t / (nbz * nby);

  const Layout &L = a.L;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * TZ + tz;
  const int n2 = L.n[2];
  constexpr bool wxy = WXY;
  const int ylo = (wxy ? min(a.lo[1] + by * (TY - 4), a.ys_max) : a.lo[1] + by * (TY - 4)) - 2;
  const int zlo = (a.zb ? min(a.lo[2] + bz * (TZ - 4), a.zs_max) : a.lo[2] + bz * (TZ - 4)) - 2;
  const int y = ylo + ty, z = zlo + tz;
  const int x0 = a.lo[0] + bxc * a.chunk;
  const int x1 = min(x0 + a.chunk - 1, a.hi[0]);
  const int nsteps = x1 - x0 + 5;   // psi^k planes x0-2 .. x1+2
  // loads: y >= lo-2 >= 0 (wxy: rows are read at their periodic images,
  // like the cyclic z columns below); on the far side rows and columns past
  // the tile's last output cell are clamped to the first halo row (cell n),
  // which the FUSE0 donor needs for psi^1 at cell n-1.  z-boundary tiles also hold
  // the halo columns.  Open (zb = 1): scalars (psi^k, psi^n) are read at the
  // open boundary's edge copy (bcond/open_3d.hpp:32-42, clamp to 0..n-1),
  // the advector from its halo faces in memory (clamped to -hw..n).
  // Cyclic (zb = 2): every field is read at the periodic image of the
  // column in 0..n-1 (z < 0: z + n-1, z > n-1: z - (n-1)), the value the
  // cyclic halo fill copies there (bcond/cyclic_3d.hpp:29-47,120-138); column n-1
  // itself holds the image of column 0.  (Columns past n+1 occur only in a
  // tile wider than the domain and feed no output: clamped to n-1.)
  const bool seam = a.zb == 2;
  auto zwrap = [n2](int c) { return c < 0 ? c + (n2 - 1) : (c > n2 - 1 ? min(c - (n2 - 1), n2 - 1) : c); };
  const int n1 = L.n[1];
  auto ywrap = [n1](int c) { return c < 0 ? c + (n1 - 1) : (c > n1 - 1 ? min(c - (n1 - 1), n1 - 1) : c); };
  const int yl = wxy ? ywrap(y) : min(y, n1);
  const int zl = seam ? zwrap(z) : (a.zb ? min(max(z, -L.hw), n2) : min(z, n2));
  const int zls = seam ? zwrap(z) : (a.zb ? min(max(z, 0), n2 - 1) : min(z, n2));
  // row and column the per-thread byte offsets count from (all offsets >= 0)
  const int zbase = seam ? -L.hw : zlo;
  const int ybase = wxy ? -L.hw : ylo;
  const long sx = L.st[0];
  const long base0 = L.org + (long)(x0 - 2) * sx + (long)ybase * L.st[1] + zbase;
  const unsigned toff = (unsigned)(((long)(yl - ybase) * L.st[1] + (zl - zbase)) * (long)sizeof(R));
  const unsigned toff_s = (unsigned)(((long)(yl - ybase) * L.st[1] + (zls - zbase)) * (long)sizeof(R));
  const bool out_ok = ty >= 2 && ty < TY - 2 && tz >= 2 && tz < TZ - 2 && y <= a.hi[1] && z <= a.hi[2];
  const bool v_ok = out_ok && y >= a.vlo[1] && y <= a.vhi[1] && z >= a.vlo[2] && z <= a.vhi[2];
  // open z-boundary tiles: block flag (uniform), halo columns (z < 0 or
  // z > n-1), the offset from a halo column to the edge column it copies,
  // and the velocity of face -1/2 (cell -1), which belongs to the domain's
  // velocity
  const bool blk_zb = a.zb == 1 && (zlo < 0 || zlo + TZ - 1 > n2 - 1);
  const bool zh = a.zb == 1 && (z < 0 || z > n2 - 1);
  const int e_off = z < 0 ? -z : (z > n2 - 1 ? n2 - 1 - z : 0);
  // lane offsets of the zero-divergence inputs of halo cells -1 and n: the
  // edge column (V_x, V_y) and the V_z face (face 1/2 of cell 0, face
  // n-3/2 of cell n-2)
  const int zx1 = z == -1 ? 1 : (z == n2 ? -1 : 0), zx2 = z == -1 ? 1 : (z == n2 ? -2 : 0);
  const bool vzm_ok = a.zb == 1 && z == -1 && ty >= 2 && ty < TY - 2 && y <= a.hi[1] && y >= a.vlo[1] &&
                      y <= a.vhi[1];
  // steps whose velocity plane (p-1) is stored: planes x0..x1 within vlo..vhi
  const int vq0 = max(3, a.vlo[0] - x0 + 3), vq1 = min(nsteps - 2, a.vhi[0] - x0 + 3);

  // running byte addresses of the tile corner at the plane each stream reads
  // next (plane 0 = x0-2); V stores at plane q-1, psi^{k+1} at plane q-2
  // typed (global) pointers with a uniform base and a 32-bit per-thread byte
  // offset: the loads use the scalar-base addressing mode, the running
  // pointers advance with scalar adds
  const long sxb = sx * (long)sizeof(R);
  auto at = [&](const R *arr, int plane) -> char * {
    return reinterpret_cast<char *>(const_cast<R *>(arr + base0)) + plane * sxb;
  };
  // load streams (wxy): plane x is read at its periodic image in 0..n0-1,
  // so a stream steps back n0-2 planes where the next plane is x = 0 or
  // x = n0 (the only crossings of a chunk's planes x0-3 .. x1+2)
  const int n0 = L.n[0];
  auto atl = [&](const R *arr, int plane) -> char * {
    const int X = x0 - 2 + plane;
    const long c = !wxy ? 0 : (X < 0 ? (long)(n0 - 1) * sxb : (X > n0 - 1 ? -(long)(n0 - 1) * sxb : 0));
    return at(arr, plane) + c;
  };
  auto dstep = [&](int X) -> long {   // stride to plane X from plane X - 1
    return wxy && (X == 0 || X == n0) ? sxb - (long)(n0 - 1) * sxb : sxb;
  };
  constexpr int AHEAD = FUSE0 ? 1 : 0;   // plane of the first "next" register set
  char *g_ps = FUSE0 ? nullptr : atl(a.psi, 0), *g_pn = FCT || FUSE0 ? atl(a.psin, AHEAD) : nullptr;
  char *g_ux = atl(a.U0, AHEAD), *g_uy = atl(a.U1, AHEAD), *g_uz = atl(a.U2, AHEAD);
  char *g_out = at(a.out, -2);
  char *g_v0 = STOREV ? at(a.V0, -1) : nullptr, *g_v1 = STOREV ? at(a.V1, -1) : nullptr,
       *g_v2 = STOREV ? at(a.V2, -1) : nullptr;
  // buffer accesses where the scalar registers suffice; the FUSE0 kernels
  // keep eight running streams live and would spill their descriptors
  auto LD = [&](const char *g) -> R {
    if constexpr (FUSE0) return *reinterpret_cast<const R *>(g + toff);
    else return buf_ld<R>(g, toff);
  };
  auto LDs = [&](const char *g) -> R {   // scalar fields (psi^k, psi^n)
    if constexpr (FUSE0) return *reinterpret_cast<const R *>(g + toff_s);
    else return buf_ld<R>(g, toff_s);
  };
  auto ST = [&](char *g, R v) {
    if constexpr (FUSE0) *reinterpret_cast<R *>(g + toff) = v;
    else buf_st<R>(g, toff, v);
  };

  // zero the frame of every LDS plane
  for (int i = tid; i < IL_N * 2 * (PW + TY); i += TY * TZ) {
    const int pl = i / (2 * (PW + TY)), r = i % (2 * (PW + TY));
    int c;
    if (r < 2 * PW) c = (r < PW ? 0 : (TY + 1) * PW) + r % PW;
    else c = (1 + (r - 2 * PW) / 2) * PW + ((r - 2 * PW) % 2 ? PW - 1 : 0);
    interior_lds<R, LN>[pl * PL + c] = R(0);
  }
  __syncthreads();
  const int ti = (ty + 1) * PW + tz + 1;
  auto S = [&](int plane, int dy, int dz) -> R & { return interior_lds<R, LN>[plane * PL + ti + dy * PW + dz]; };

  // FUSE0 frame job of this thread: one psi^n, GC_y or GC_z value outside
  // the tile per plane (rows y = -1 and Y, columns z = -1 and Z of psi^n;
  // row y = -1 of GC_y; column z = -1 of GC_z).  Threads without a job
  // reload their own psi^n cell into a spare LDS slot, so the job needs no
  // branch.
  R fr_sc = 1;
  int fr_lds = IL_N * PL + tid;
  const char *fr_g = atl(a.psin, 0) + toff_s;
  if (FUSE0 && tid < ITile<TYV>::NFRAME) {
    int fy, fz, kind;
    if (tid < 2 * TZ) { kind = 0; fy = tid < TZ ? -1 : TY; fz = tid % TZ; }
    else if (tid < 2 * TZ + 2 * TY) { kind = 0; const int r = tid - 2 * TZ; fy = r % TY; fz = r < TY ? -1 : TZ; }
    else if (tid < 3 * TZ + 2 * TY) { kind = 1; fy = -1; fz = tid - 2 * TZ - 2 * TY; }
    else { kind = 2; fy = tid - 3 * TZ - 2 * TY; fz = -1; }
    const int gy = wxy ? ywrap(ylo + fy) : min(ylo + fy, n1);
    const int gz = seam ? zwrap(zlo + fz)
                        : (!a.zb ? min(zlo + fz, n2) : (kind == 0 ? min(max(zlo + fz, 0), n2 - 1)
                                                                  : min(max(zlo + fz, -L.hw), n2)));
    const long off = ((long)(gy - ybase) * L.st[1] + (gz - zbase)) * (long)sizeof(R);
    const R *arr = kind == 0 ? a.psin : (kind == 1 ? a.U1 : a.U2);
    fr_g = atl(arr, 0) + off;
    fr_lds = (kind == 0 ? IL_PN : (kind == 1 ? IL_UY : IL_UZ)) * PL + (fy + 1) * PW + fz + 1;
    fr_sc = kind == 0 ? a.in_scale : R(1);
  }
  auto FLD = [&](int plane) -> R {
    return *reinterpret_cast<const R *>(fr_g + plane * sxb);
  };

  // register rings (slot = plane index mod 6)
  R ps[6], pn[6], ux[6], uy[6], uz[6];
  R vx_[2] = {0, 0}, fpx_[2] = {0, 0}, bu_[2] = {0, 0}, bd_[2] = {0, 0}, fx_[2] = {0, 0};
  for (int i = 0; i < 6; ++i) ps[i] = pn[i] = ux[i] = uy[i] = uz[i] = R(0);
  R n_ps = 0, n_pn = 0, n_ux = 0, n_uy = 0, n_uz = 0, n_fr = 0;
  const R sc = a.in_scale;

  if (!FUSE0) {
    n_ps = LDs(g_ps);
    if (FCT) n_pn = LDs(g_pn);
    n_ux = LD(g_ux); n_uy = LD(g_uy); n_uz = LD(g_uz);
  } else {
    // prologue: psi^1 on plane 0 from psi^n planes -1..1 and GC planes -1, 0
    pn[5] = LDs(atl(a.psin, -1)) * sc;
    pn[0] = LDs(atl(a.psin, 0)) * sc;
    pn[1] = LDs(atl(a.psin, 1)) * sc;
    ux[5] = LD(atl(a.U0, -1));
    ux[0] = LD(atl(a.U0, 0));
    uy[0] = LD(atl(a.U1, 0));
    uz[0] = LD(atl(a.U2, 0));
    S(IL_PN + 0, 0, 0) = pn[0];
    S(IL_UX + 0, 0, 0) = ux[0];
    S(IL_UY + 0, 0, 0) = uy[0];
    S(IL_UZ + 0, 0, 0) = uz[0];
    interior_lds<R, LN>[fr_lds] = FLD(0) * fr_sc;
    // plane 1 for the first step's publish
    n_pn = LDs(g_pn); n_ux = LD(g_ux); n_uy = LD(g_uy); n_uz = LD(g_uz);
    fr_g += dstep(x0 - 1);
    n_fr = *reinterpret_cast<const R *>(fr_g);
    __syncthreads();
    const R pc = pn[0];
    const R f0 = iflux(pc, pn[1], ux[0]), f1 = iflux(pn[5], pc, ux[5]);
    const R f2 = iflux(pc, S(IL_PN, 1, 0), uy[0]), f3 = iflux(S(IL_PN, -1, 0), pc, S(IL_UY, -1, 0));
    const R f4 = iflux(pc, S(IL_PN, 0, 1), uz[0]), f5 = iflux(S(IL_PN, 0, -1), pc, S(IL_UZ, 0, -1));
    ps[0] = pc + ((-f0 + f1) + (-f2 + f3) + (-f4 + f5)) / 1;
    if (blk_zb) {
      // halo columns of psi^1: the open boundary's edge copy, from the
      // edge column's lane (a tile row lies in one wave)
      const R v = __shfl(ps[0], (int)__lane_id() + e_off);
      if (zh) ps[0] = v;
    }
    S(IL_PSI + 0, 0, 0) = ps[0];
  }

  auto step = [&](auto PHc, const int q) __attribute__((always_inline)) {
    constexpr int PH = decltype(PHc)::value;
    // ring slots of plane p + k
    auto s6 = [](int k) constexpr { return ((PH + k) % 6 + 6) % 6; };
    auto s3 = [](int k) constexpr { return ((PH + k) % 3 + 3) % 3; };
    auto s2 = [](int k) constexpr { return ((PH + k) % 2 + 2) % 2; };
    // uy / uz / psi^n LDS rings: 2 planes, 3 with FUSE0
    auto su = [](int k) constexpr { return FUSE0 ? ((PH + k) % 3 + 3) % 3 : ((PH + k) % 2 + 2) % 2; };

    // ---- stage 1: publish plane p (p+1 with FUSE0), prefetch the next one
    if (!FUSE0) {
      ps[s6(0)] = n_ps; ux[s6(0)] = n_ux; uy[s6(0)] = n_uy; uz[s6(0)] = n_uz;
      S(IL_PSI + s3(0), 0, 0) = n_ps;
      S(IL_UX + s3(0), 0, 0) = n_ux;
      S(IL_UY + su(0), 0, 0) = n_uy;
      S(IL_UZ + su(0), 0, 0) = n_uz;
      if (FCT) {
        pn[s6(0)] = n_pn * sc;
        S(IL_PN + su(0), 0, 0) = pn[s6(0)];
      }
      if (q + 1 < nsteps) {
        const long ds = dstep(x0 - 1 + q);   // to plane q + 1
        g_ps += ds; g_ux += ds; g_uy += ds; g_uz += ds;
        n_ps = LDs(g_ps);
        if (FCT) { g_pn += ds; n_pn = LDs(g_pn); }
        n_ux = LD(g_ux); n_uy = LD(g_uy); n_uz = LD(g_uz);
      }
    } else {
      ux[s6(1)] = n_ux; uy[s6(1)] = n_uy; uz[s6(1)] = n_uz;
      pn[s6(1)] = n_pn * sc;
      S(IL_UY + su(1), 0, 0) = n_uy;
      S(IL_UZ + su(1), 0, 0) = n_uz;
      S(IL_PN + su(1), 0, 0) = pn[s6(1)];
      interior_lds<R, LN>[fr_lds + (fr_lds < IL_N * PL ? su(1) * PL : 0)] = n_fr * fr_sc;
      const long ds = dstep(x0 + q);   // to plane q + 2
      if (q + 1 < nsteps) { g_pn += ds; n_pn = LDs(g_pn); }
      if (q + 2 < nsteps) {
        g_ux += ds; g_uy += ds; g_uz += ds;
        n_ux = LD(g_ux); n_uy = LD(g_uy); n_uz = LD(g_uz);
        fr_g += ds;
        n_fr = *reinterpret_cast<const R *>(fr_g);
      }
    }
    __syncthreads();

    // ---- stage 2: velocity on plane p-1 (V_x on the face owned by p-1)
    R vx = 0, vy = 0, vz = 0;
    if (q >= 1) {
      {
        // V_x at face x+1/2: normal x (planes p-1, p), tangential (y, z)
        auto P = [&](int da, int db, int dc) -> R {
          if (db == 0 && dc == 0) return da == 0 ? ps[s6(-1)] : ps[s6(0)];
          return da == 0 ? S(IL_PSI + s3(-1), db, dc) : S(IL_PSI + s3(0), db, dc);
        };
        const R s1 = (uy[s6(0)] + uy[s6(-1)] + S(IL_UY + su(0), -1, 0) + S(IL_UY + su(-1), -1, 0));
        const R s2 = (uz[s6(0)] + uz[s6(-1)] + S(IL_UZ + su(0), 0, -1) + S(IL_UZ + su(-1), 0, -1));
        vx = antidiff_core<R, G, 3, IGA, false, false, decltype(P), true>(ux[s6(-1)], s1, s2, P);
      }
      // halo columns at an open z boundary: the normal-component fill sets
      // V_x and V_y there to zero (bcond/open_3d.hpp fill_halos_vctr_nrml)
      if (zh) vx = R(0);
      if (q >= 2) {
        {
          // V_y at face y+1/2 of plane p-1: normal y, tangential (z, x)
          auto P = [&](int da, int db, int dc) -> R {
            if (da == 0 && db == 0) return ps[s6(dc - 1)];
            return S(IL_PSI + s3(dc - 1), da, db);
          };
          const R s1 = (S(IL_UZ + su(-1), 1, 0) + uz[s6(-1)] + S(IL_UZ + su(-1), 1, -1) + S(IL_UZ + su(-1), 0, -1));
          const R s2 = (S(IL_UX + s3(-1), 1, 0) + ux[s6(-1)] + S(IL_UX + s3(-2), 1, 0) + ux[s6(-2)]);
          vy = antidiff_core<R, G, 3, IGA, false, false, decltype(P), true>(uy[s6(-1)], s1, s2, P);
        }
        {
          // V_z at face z+1/2 of plane p-1: normal z, tangential (x, y)
          auto P = [&](int da, int db, int dc) -> R {
            if (da == 0 && dc == 0) return ps[s6(db - 1)];
            return S(IL_PSI + s3(db - 1), dc, da);
          };
          const R s1 = (S(IL_UX + s3(-1), 0, 1) + ux[s6(-1)] + S(IL_UX + s3(-2), 0, 1) + ux[s6(-2)]);
          const R s2 = (S(IL_UY + su(-1), 0, 1) + uy[s6(-1)] + S(IL_UY + su(-1), -1, 1) + S(IL_UY + su(-1), -1, 0));
          vz = antidiff_core<R, G, 3, IGA, false, false, decltype(P), true>(uz[s6(-1)], s1, s2, P);
        }
        if (zh) vy = R(0);
        if (FCT) {
          S(IL_VY, 0, 0) = vy;
          S(IL_VZ, 0, 0) = vz;
        }
      }
      if (STOREV && q >= vq0 && q <= vq1) {
        if (v_ok) {
          ST(g_v0, vx);
          ST(g_v1, vy);
        }
        if (v_ok || vzm_ok) ST(g_v2, vz);
      }
    }
    if (STOREV) { g_v0 += sxb; g_v1 += sxb; g_v2 += sxb; }
    __syncthreads();

    // ---- stage 3: beta on plane p-1 (FCT)
    R bu = 0, bd = 0;
    const R fpx = IGA ? vx : iflux(ps[s6(-1)], ps[s6(0)], vx);   // unlimited flux, face owned by p-1
    R yp = 0, zp = 0;
    if (FCT && q >= 2) {
      const R pc = ps[s6(-1)];
      yp = S(IL_PSI + s3(-1), 1, 0);
      zp = S(IL_PSI + s3(-1), 0, 1);
      const R ym = S(IL_PSI + s3(-1), -1, 0), zm = S(IL_PSI + s3(-1), 0, -1);
      const int pm = IL_PN + su(-1);
      const R pmax = fmx(fmx(fmx(pn[s6(-1)], pn[s6(0)]), fmx(pn[s6(-2)], S(pm, 1, 0))),
                         fmx(S(pm, -1, 0), fmx(S(pm, 0, 1), S(pm, 0, -1))));
      const R pmin = fmn(fmn(fmn(pn[s6(-1)], pn[s6(0)]), fmn(pn[s6(-2)], S(pm, 1, 0))),
                         fmn(S(pm, -1, 0), fmn(S(pm, 0, 1), S(pm, 0, -1))));
      const R vym = S(IL_VY, -1, 0);
      R vzm = S(IL_VZ, 0, -1), vzp = vz;
      if (blk_zb) {
        // V_z on the outer halo face next to halo cell -1 (face -3/2) and
        // halo cell n (face n+1/2) of an open z boundary: the zero-divergence
        // fill (bcond/open_3d.hpp fill_halos_vctr_alng) from the edge cell's
        // velocities, computed by the only thread that reads it (beta of
        // that halo cell).  V_x of the edge column comes from its lane (a
        // tile row lies in one wave), V_y from the published plane.
        const int ln = (int)__lane_id();
        const R ex = __shfl(vx, ln + zx1), exp_ = __shfl(vx_[s2(-1)], ln + zx1), ez = __shfl(vz, ln + zx2);
        const R eyp = interior_lds<R, LN>[IL_VY * PL + ti + zx1], eym = interior_lds<R, LN>[IL_VY * PL + ti - PW + zx1];
        if (z == -1) vzm = ez - (exp_ - ex) - (eym - eyp);
        if (z == n2) vzp = ez + (exp_ - ex) + (eym - eyp);
      }
      const R fyp = IGA ? vy : iflux(pc, yp, vy);
      const R fym = IGA ? vym : iflux(ym, pc, vym);
      const R fzp = IGA ? vzp : iflux(pc, zp, vzp);
      const R fzm = IGA ? vzm : iflux(zm, pc, vzm);
      ibeta<R, G>(bu, bd, pc, ps[s6(0)], ps[s6(-2)], yp, ym, zp, zm, pmin, pmax, fpx, fpx_[s2(-1)],
                  fyp, fym, fzp, fzm);
      S(IL_BU, 0, 0) = bu;
      S(IL_BD, 0, 0) = bd;
    }
    // FCT: the x flux of the face owned by p-2 (between planes p-2 and p-1,
    // limited with beta of both planes, in registers) and psi^{k+1} on
    // plane p-2, whose y and z fluxes the previous step published; stage 4
    // overwrites them only after the barrier
    R fx = 0;
    if (FCT && q >= 3) {
      const R mx = imono(vx_[s2(-1)], bu_[s2(-1)], bd_[s2(-1)], bu, bd);
      fx = IGA ? mx : iflux_v(ps[s6(-2)], ps[s6(-1)], mx, vx_[s2(-1)]);
      if (q >= 4 && out_ok) {
        const R r = ps[s6(-2)] + ((-fx + fx_[s2(-1)]) + (-S(IL_FY, 0, 0) + S(IL_FY, -1, 0))
                                  + (-S(IL_FZ, 0, 0) + S(IL_FZ, 0, -1))) / 1;
        ST(g_out, r * a.out_scale);
      }
    }
    if (FCT) __syncthreads();

    // ---- stage 4: limited y and z fluxes of plane p-1; without FCT also
    // the x flux of the face owned by p-2 and psi^{k+1} on plane p-2
    if (q >= 3) {
      if (!FCT) {
        fx = IGA ? vx_[s2(-1)] : iflux_v(ps[s6(-2)], ps[s6(-1)], vx_[s2(-1)], vx_[s2(-1)]);
        yp = S(IL_PSI + s3(-1), 1, 0);
        zp = S(IL_PSI + s3(-1), 0, 1);
      }
      // flux planes: FCT one each; without FCT a ring, plane p-1 in slot
      // s2(-1) and plane p-2 in s2(-2)
      const int wf = FCT ? 0 : s2(-1);
      const R my = FCT ? imono(vy, bu, bd, S(IL_BU, 1, 0), S(IL_BD, 1, 0)) : vy;
      const R mz = FCT ? imono(vz, bu, bd, S(IL_BU, 0, 1), S(IL_BD, 0, 1)) : vz;
      S(IL_FY + wf, 0, 0) = IGA ? my : iflux_v(ps[s6(-1)], yp, my, vy);
      S(IL_FZ + wf, 0, 0) = IGA ? mz : iflux_v(ps[s6(-1)], zp, mz, vz);
      if (!FCT && q >= 4 && out_ok) {
        const R r = ps[s6(-2)] + ((-fx + fx_[s2(-1)]) + (-S(IL_FY + s2(-2), 0, 0) + S(IL_FY + s2(-2), -1, 0))
                                  + (-S(IL_FZ + s2(-2), 0, 0) + S(IL_FZ + s2(-2), 0, -1))) / 1;
        ST(g_out, r * a.out_scale);
      }
      fx_[s2(0)] = fx;
    }
    g_out += sxb;
    vx_[s2(0)] = vx;
    fpx_[s2(0)] = fpx;
    bu_[s2(0)] = bu;
    bd_[s2(0)] = bd;

    // ---- FUSE0: psi^1 on plane p+1 (donor cell from psi^n, GC)
    if (FUSE0) {
      S(IL_UX + s3(1), 0, 0) = ux[s6(1)];
      if (q + 1 < nsteps) {
        const int pp = IL_PN + su(1);
        const R pc = pn[s6(1)];
        const R f0 = iflux(pc, n_pn * sc, ux[s6(1)]), f1 = iflux(pn[s6(0)], pc, ux[s6(0)]);
        const R f2 = iflux(pc, S(pp, 1, 0), uy[s6(1)]), f3 = iflux(S(pp, -1, 0), pc, S(IL_UY + su(1), -1, 0));
        const R f4 = iflux(pc, S(pp, 0, 1), uz[s6(1)]), f5 = iflux(S(pp, 0, -1), pc, S(IL_UZ + su(1), 0, -1));
        ps[s6(1)] = pc + ((-f0 + f1) + (-f2 + f3) + (-f4 + f5)) / 1;
        if (blk_zb) {
          const R v = __shfl(ps[s6(1)], (int)__lane_id() + e_off);
          if (zh) ps[s6(1)] = v;
        }
        S(IL_PSI + s3(1), 0, 0) = ps[s6(1)];
      }
    }
  };

  // no peeling of the first iteration: a peeled copy may round differently
  // (other contraction choices), which would make the results depend on
  // where the x chunks start
#pragma clang loop unroll(disable)
  for (int q = 0; q < nsteps; q += 6) {
    step(ic<0>(), q);
    if (q + 1 >= nsteps) break;
    step(ic<1>(), q + 1);
    if (q + 2 >= nsteps) break;
    step(ic<2>(), q + 2);
    if (q + 3 >= nsteps) break;
    step(ic<3>(), q + 3);
    if (q + 4 >= nsteps) break;
    step(ic<4>(), q + 4);
    if (q + 5 >= nsteps) break;
    step(ic<5>(), q + 5);
  }
}

template <class R, class G, bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV, bool WXY>
__global__ void __launch_bounds__(ITile<TYV>::Y * ITile<TYV>::Z) MPDX_INTERIOR_ATTR
k_fz_interior(InteriorArgs<R> a)
{
  fz_interior_block<R, G, FCT, IGA, STOREV, FUSE0, TYV, WXY>(a, blockIdx.x);
}

// ---------------------------------------------------------------- host side

inline bool fused_supported(const mpdx_params &, const Layout &)
{
  // every option: the interior and 2D pass kernels take the common ones,
  // the boundary-shell kernels (over the whole grid) the rest, khn included
  return true;
}

// the interior kernel covers 3D, standard or iga antidiff without abs/npa/pfc
inline bool interior_supported(const mpdx_params &p, const Layout &L)
{
  if (L.nd != 3) return false;
  if (p.opts & (O_ABS | O_NPA | O_PFC | O_KHN)) return false;
  for (int d = 0; d < 3; ++d)
    if (L.n[d] < 12) return false;
  return true;
}

inline BcMap make_bcmap(const mpdx_solver *s)
{
  BcMap m;
  for (int d = 0; d < 3; ++d) {
    m.n[d] = s->L.n[d];
    for (int sd = 0; sd < 2; ++sd) m.cyc[d][sd] = d < s->L.nd && s->p.bc[d][sd] == MPDX_BC_CYCLIC;
  }
  return m;
}

}  // namespace mpdx

// Per-pass MPDATA kernels: one kernel per reference pass over a full
// halo-extended array, plus small kernels that replay the reference's halo
// fills with the reference's ranges and in the reference's order.  This path
// is exact for every input (including cyclic fields whose duplicated edge
// cells disagree) and serves every option combination; the fused kernels
// (mpdx_fused.hpp) are checked against it.
//
// Citations are path:line in the reference tree (libmpdata++/...).
#pragma once

#include <type_traits>

#include "mpdx_common.hpp"

namespace mpdx {

template <class R> __device__ inline R pos_rt(unsigned o, R x)
{
  return (o & O_NPA) ? pospart<true>(x) : pospart<false>(x);
}
template <class R> __device__ inline R neg_rt(unsigned o, R x)
{
  return (o & O_NPA) ? negpart<true>(x) : negpart<false>(x);
}
template <class G, class R> __device__ inline R frac_rt(unsigned o, R n, R d)
{
  return (o & O_PFC) ? frac<true, G>(n, d) : frac<false, G>(n, d);
}

// ---------------------------------------------------------------- halos

// scalar / tangential-advector halo along d over the rotated (b, c, m) box:
// mode 0 cyclic copy with period N-1 (bcond/cyclic_3d.hpp:29-33,120-124),
// mode 1 edge copy (bcond/open_3d.hpp:32-42,154-164), mode 2 zero
// (open tangential advector, open_3d.hpp:127-133,248-254)
template <class R>
__device__ inline void fill_one(R *a, const Rot &r, int b, int c, int m, int side, int mode, int N)
{
  long dst, src;
  if (side == 0) { dst = rr(r, -m, b, c); src = rr(r, mode == 0 ? N - 1 - m : 0, b, c); }
  else           { dst = rr(r, N - 1 + m, b, c); src = rr(r, mode == 0 ? m : N - 1, b, c); }
  a[dst] = mode == 2 ? R(0) : a[src];
}

// up to 6 fills that touch disjoint cells and read none of each other's
// outputs (both sides of one fill; fills of different arrays), one per
// grid z index.  The box order puts the memory-contiguous coordinate last
// (the fastest thread index): ord 0 is (b, m, c), ord 1 (2D) is (m, c, b).
template <class R> struct FillBatch {
  R *a[6];
  Rot r[6];
  Box bx[6];
  int side[6], mode[6], N[6], ord[6];
  int n;
};

template <class R>
__global__ void k_fill_batch(FillBatch<R> fb)
{
  const int q = blockIdx.z;
  if ((int)blockIdx.y >= fb.bx[q].ext[0]) return;
  int b, c, m;
  if (fb.ord[q] == 0) {
    if (!box_idx(fb.bx[q], b, m, c)) return;
  } else if (!box_idx(fb.bx[q], m, c, b)) {
    return;
  }
  fill_one(fb.a[q], fb.r[q], b, c, m, fb.side[q], fb.mode[q], fb.N[q]);
}

// open-boundary zeroing of perpendicular advector components at the edge
// (open_3d.hpp:79-103,201-225; open_2d.hpp fill_halos_vctr_alng)
template <class R>
__device__ __forceinline__ void alng_zero_one(R *av1, R *av2, const Rot &r, int nd, int d, int e, int raf,
                                              int ral, int rbf, int rbl, int t)
{
  if (nd == 3) {
    if (d == 0 && t <= ral - raf) {
      int b = raf + t;
      av2[rr(r, e, b, rbf - 1)] = R(0);
      av2[rr(r, e, b, rbl)] = R(0);
    }
    if ((d == 0 || d == 1) && t <= rbl - rbf) {
      int c = rbf + t;
      av1[rr(r, e, raf - 1, c)] = R(0);
      av1[rr(r, e, ral, c)] = R(0);
    }
  } else if (d == 0 && t == 0) {
    av1[rr(r, e, raf - 1, 0)] = R(0);
    av1[rr(r, e, ral, 0)] = R(0);
  }
}

template <class R>
__global__ void k_alng_zero(R *av1, R *av2, Rot r, int nd, int d, int e0, int e1,
                            int raf, int ral, int rbf, int rbl)
{
  // grid y: the edge e0 (first side) or e1 (second side); the two touch
  // different cells
  const int e = blockIdx.y == 0 ? e0 : e1;
  alng_zero_one(av1, av2, r, nd, d, e, raf, ral, rbf, rbl, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// normal advector halo faces: cyclic copy (cyclic_3d.hpp:44-48,135-139) or
// the open zero-divergence condition (open_3d.hpp:111-124,233-245); the box
// is (ii or m, b, c).  av1 varies along b, av2 along c; cfirst says that the
// reference's first difference term (component d+1) is the one along c (the
// host swaps the tangential axes to make c the contiguous one), so the sum
// keeps the reference's order.
template <class R>
__device__ inline void alng_one(R *avd, const R *av1, const R *av2, const Rot &r, int ii, int b, int c,
                                int nd, int side, int cyclic, int N, int cfirst)
{
  if (cyclic) {
    const int m = ii;
    if (side == 0) avd[rr(r, -m, b, c)] = avd[rr(r, N - 1 - m, b, c)];
    else           avd[rr(r, N - 2 + m, b, c)] = avd[rr(r, m - 1, b, c)];
    return;
  }
  const int e = side == 0 ? 0 : N - 1;
  R v;
  if (nd == 3) {
    const R tb = av1[rr(r, e, b - 1, c)] - av1[rr(r, e, b, c)];
    const R tc = av2[rr(r, e, b, c - 1)] - av2[rr(r, e, b, c)];
    const R t1 = cfirst ? tc : tb, t2 = cfirst ? tb : tc;
    if (side == 0)
      v = avd[rr(r, e, b, c)] - t1 - t2;
    else
      v = avd[rr(r, e - 1, b, c)] + t1 + t2;
  } else {
    if (side == 0)
      v = avd[rr(r, e, b, 0)] - (av1[rr(r, e, b - 1, 0)] - av1[rr(r, e, b, 0)]);
    else
      v = avd[rr(r, e - 1, b, 0)] + (av1[rr(r, e, b - 1, 0)] - av1[rr(r, e, b, 0)]);
  }
  avd[rr(r, ii, b, c)] = v;
}

// both sides of one dimension's normal-component fill in one launch (grid z =
// side): each side writes its own halo faces and reads only cells the other
// side does not write
template <class R>
__global__ void k_alng2(R *avd, const R *av1, const R *av2, Rot r, Box bx0, Box bx1, int nd,
                        int cyc0, int cyc1, int N, int cfirst)
{
  const int side = blockIdx.z;
  const Box &bx = side == 0 ? bx0 : bx1;
  if ((int)blockIdx.y >= bx.ext[0]) return;
  int ii, b, c;
  if (!box_idx(bx, ii, b, c)) return;
  alng_one(avd, av1, av2, r, ii, b, c, nd, side, side == 0 ? cyc0 : cyc1, N, cfirst);
}

// one dimension's normal-component fill over a box list (box tag = side),
// for fills restricted to some rows of the halo plane
template <class R>
__global__ void k_alng_list(R *avd, const R *av1, const R *av2, Rot r, BoxList bl, int nd, int cyc0,
                            int cyc1, int N, int cfirst)
{
  int q, ii, b, c;
  if (!boxlist_idx(bl, q, ii, b, c)) return;
  const int side = bl.tag[q];
  alng_one(avd, av1, av2, r, ii, b, c, nd, side, side == 0 ? cyc0 : cyc1, N, cfirst);
}

// ---------------------------------------------------------------- passes

// donor-cell fluxes along d (donorcell_formulae.hpp:50-79)
template <class R>
__global__ void k_flux(R *out, const R *psi, const R *C, Rot r, Box bx, unsigned o)
{
  int a, b, c;
  if (!box_idx(bx, a, b, c)) return;
  const long p0 = rr(r, a, b, c);
  const R cc = C[p0];
  out[p0] = pos_rt(o, cc) * psi[p0] + neg_rt(o, cc) * psi[rr(r, a + 1, b, c)];
}

// donorcell_sum over the interior (donorcell_formulae.hpp:105-165)
template <class R>
__global__ void k_donorcell(R *psi_new, const R *psi_old, const R *f0, const R *f1,
                            const R *f2, Layout L, Box bx, unsigned o)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  const long p = L.org + i * L.st[0] + j * L.st[1] + k * L.st[2];
  R f[6];
  f[0] = f0[p]; f[1] = f0[p - L.st[0]];
  f[2] = f1[p]; f[3] = f1[p - L.st[1]];
  if (L.nd == 3) { f[4] = f2[p]; f[5] = f2[p - L.st[2]]; }
  if (!(o & O_KHN)) {
    if (L.nd == 3)
      psi_new[p] = psi_old[p] + ((-f[0] + f[1]) + (-f[2] + f[3]) + (-f[4] + f[5])) / 1;
    else
      psi_new[p] = psi_old[p] + ((-f[0] + f[1]) + (-f[2] + f[3])) / 1;
  } else {
    // kahan_zro / kahan_add sequence (formulae/kahan_sum.hpp:13-27)
    R sum = 0, cc = 0;
    R in[7];
    in[0] = psi_old[p];
    const int nq = 2 * L.nd;
    for (int q = 0; q < nq; ++q) in[1 + q] = (q % 2 == 0) ? -f[q] / 1 : f[q] / 1;
    for (int q = 0; q <= nq; ++q) {
      R y = in[q] - cc;
      R tt = sum + y;
      cc = (tt - sum) - y;
      sum = tt;
    }
    psi_new[p] = sum;
  }
}

// antidiffusive velocity along d, standard form: formulae_mpdata_3d.hpp:253-295
// (2D: formulae_mpdata_2d.hpp:241-278) over faces -1..N_d-1 x interior
template <class R, class G>
__global__ void k_antidiff(R *V, const R *psi, const R *U0, const R *U1, const R *U2,
                           Rot r, Box bx, int nd, unsigned o)
{
  int a, b, c;
  if (!box_idx(bx, a, b, c)) return;
  const bool iga = o & O_IGA, ab = o & O_ABS;
  auto P = [&](int x, int y, int z) -> R {
    R v = psi[rr(r, x, y, z)];
    return ab ? fabs(v) : v;
  };
  const long p0 = rr(r, a, b, c);
  const R u = U0[p0];
  R ndx, ndy;
  if (iga) {
    ndx = nd == 3 ? psi[rr(r, a + 1, b, c)] - psi[p0]
                  : 2 * (psi[rr(r, a + 1, b, c)] - psi[p0]) / (1 + 1);
    ndy = (psi[rr(r, a + 1, b + 1, c)] + psi[rr(r, a, b + 1, c)]
           - psi[rr(r, a + 1, b - 1, c)] - psi[rr(r, a, b - 1, c)]) / 4;
  } else {
    ndx = 2 * frac_rt<G>(o, P(a + 1, b, c) - P(a, b, c), P(a + 1, b, c) + P(a, b, c));
    ndy = frac_rt<G>(o, P(a + 1, b + 1, c) + P(a, b + 1, c) - P(a + 1, b - 1, c) - P(a, b - 1, c),
                        P(a + 1, b + 1, c) + P(a, b + 1, c) + P(a + 1, b - 1, c) + P(a, b - 1, c));
  }
  const R g1 = (U1[rr(r, a + 1, b, c)] + U1[rr(r, a, b, c)]
                + U1[rr(r, a + 1, b - 1, c)] + U1[rr(r, a, b - 1, c)]) / 4;
  if (nd == 3) {
    R ndz;
    if (iga)
      ndz = (psi[rr(r, a + 1, b, c + 1)] + psi[rr(r, a, b, c + 1)]
             - psi[rr(r, a + 1, b, c - 1)] - psi[rr(r, a, b, c - 1)]) / 4;
    else
      ndz = frac_rt<G>(o, P(a + 1, b, c + 1) + P(a, b, c + 1) - P(a + 1, b, c - 1) - P(a, b, c - 1),
                          P(a + 1, b, c + 1) + P(a, b, c + 1) + P(a + 1, b, c - 1) + P(a, b, c - 1));
    const R g2 = (U2[rr(r, a + 1, b, c)] + U2[rr(r, a, b, c)]
                  + U2[rr(r, a + 1, b, c - 1)] + U2[rr(r, a, b, c - 1)]) / 4;
    V[p0] = fabs(u) / 2 * (1 - fabs(u) / 1) * ndx - u / 2 * (g1 * ndy + g2 * ndz) / 1;
  } else {
    V[p0] = fabs(u) / 2 * (1 - fabs(u) / 1) * ndx - u * g1 / (2 * 1) * ndy;
  }
}

// psi_min / psi_max over interior^1 (mpdata_fct_3d.hpp:31-55, mpdata_fct_2d.hpp:31-46)
template <class R>
__global__ void k_fct_init(R *pmin, R *pmax, const R *psi, Layout L, Box bx)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  const long p = L.org + i * L.st[0] + j * L.st[1] + k * L.st[2];
  const long s0 = L.st[0], s1 = L.st[1], s2 = L.st[2];
  if (L.nd == 3) {
    pmin[p] = smin(smin(smin(smin(smin(smin(psi[p], psi[p + s0]), psi[p - s0]), psi[p + s1]),
                              psi[p - s1]), psi[p + s2]), psi[p - s2]);
    pmax[p] = smax(smax(smax(smax(smax(smax(psi[p], psi[p + s0]), psi[p - s0]), psi[p + s1]),
                              psi[p - s1]), psi[p + s2]), psi[p - s2]);
  } else {
    pmin[p] = smin(smin(smin(smin(psi[p + s1], psi[p - s0]), psi[p]), psi[p + s0]), psi[p - s1]);
    pmax[p] = smax(smax(smax(smax(psi[p + s1], psi[p - s0]), psi[p]), psi[p + s0]), psi[p - s1]);
  }
}

// beta_up / beta_dn over interior^1: formulae_mpdata_fct_3d.hpp:21-145,
// formulae_mpdata_fct_2d.hpp:21-120
template <class R, class G>
__global__ void k_betas(R *bup, R *bdn, const R *psi, const R *pmin, const R *pmax,
                        const R *f0, const R *f1, const R *f2, Layout L, Box bx, unsigned o)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  const long p = L.org + i * L.st[0] + j * L.st[1] + k * L.st[2];
  const long s0 = L.st[0], s1 = L.st[1], s2 = L.st[2];
  R up_nom, dn_nom, up_den, dn_den;
  if (L.nd == 3) {
    up_nom = (smax(pmax[p], smax(psi[p], smax(psi[p + s0], smax(psi[p - s0], smax(psi[p + s1],
              smax(psi[p - s1], smax(psi[p + s2], psi[p - s2]))))))) - psi[p]) * 1;
    dn_nom = (psi[p] - smin(pmin[p], smin(psi[p + s1], smin(psi[p - s0], smin(psi[p],
              smin(psi[p + s0], smin(psi[p + s2], smin(psi[p - s2], psi[p - s1])))))))) * 1;
    up_den = (pos_rt(o, f0[p - s0]) - neg_rt(o, f0[p]))
           + (pos_rt(o, f1[p - s1]) - neg_rt(o, f1[p]))
           + (pos_rt(o, f2[p - s2]) - neg_rt(o, f2[p]));
    dn_den = (pos_rt(o, f0[p]) - neg_rt(o, f0[p - s0]))
           + (pos_rt(o, f1[p]) - neg_rt(o, f1[p - s1]))
           + (pos_rt(o, f2[p]) - neg_rt(o, f2[p - s2]));
  } else {
    up_nom = (smax(pmax[p], smax(psi[p + s1], smax(psi[p - s0], smax(psi[p],
              smax(psi[p + s0], psi[p - s1]))))) - psi[p]) * 1;
    dn_nom = (psi[p] - smin(pmin[p], smin(psi[p + s1], smin(psi[p - s0], smin(psi[p],
              smin(psi[p + s0], psi[p - s1])))))) * 1;
    up_den = (pos_rt(o, f0[p - s0]) - neg_rt(o, f0[p]))
           + (pos_rt(o, f1[p - s1]) - neg_rt(o, f1[p]));
    dn_den = (pos_rt(o, f0[p]) - neg_rt(o, f0[p - s0]))
           + (pos_rt(o, f1[p]) - neg_rt(o, f1[p - s1]));
  }
  bup[p] = fct_frac<G>(up_nom, up_den);
  bdn[p] = fct_frac<G>(dn_nom, dn_den);
}

// limited antidiffusive velocity: formulae_mpdata_fct_3d.hpp:147-247
template <class R>
__global__ void k_gc_mono(R *M, const R *V, const R *psi, const R *bup, const R *bdn,
                          Rot r, Box bx, unsigned o)
{
  int a, b, c;
  if (!box_idx(bx, a, b, c)) return;
  const long p0 = rr(r, a, b, c), p1 = rr(r, a + 1, b, c);
  const R v = V[p0];
  const R m_dn_up = smin(R(1), smin(bdn[p0], bup[p1]));
  const R m_up_dn = smin(R(1), smin(bup[p0], bdn[p1]));
  R m;
  const bool absv = (o & O_ABS) && !(o & O_IGA);
  if (!absv) m = v > 0 ? m_dn_up : m_up_dn;
  else if (v > 0) m = psi[p0] > 0 ? m_dn_up : m_up_dn;
  else m = psi[p1] > 0 ? m_up_dn : m_dn_up;
  M[p0] = v * m;
}

// state scaling by an exact power of two (solver_common.hpp:374-379)
template <class R>
__global__ void k_scale(R *a, Layout L, Box bx, R factor)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  a[L.org + i * L.st[0] + j * L.st[1] + k * L.st[2]] *= factor;
}

// reversing deformational flow from a streamfunction on cell corners
// (the analytic advector of the reference's reversing_deform test family,
// tests/mp3_paper_2018_JCP/reversing_deform/reversing_deform.hpp:30-105):
// corner value psi(a, b) = Q(amp * sx2[a] * sy2[b]) with sx2[a] =
// sin^2(pi (a + 1/2)/(n0 - 1)) (sy2 likewise along y, both tabulated on the
// host from a = -2) and Q rounding to a multiple of 2^-17;
// GC_x = psi(i, j) - psi(i, j-1), GC_y = psi(i-1, j) - psi(i, j), GC_z = 0
__device__ inline double deform_corner(int a, int b, const double *sx2, const double *sy2, double amp)
{
  return rint(amp * sx2[a + 2] * sy2[b + 2] * 131072.0) / 131072.0;
}

// amp_tab[*step] is the amplitude of the current step (a table uploaded per
// advance call, the step counter advanced on the device by k_tick), so a
// captured time step can be replayed unchanged
template <class R>
__global__ void k_gen_deform(R *g0, R *g1, R *g2, Layout L, Box bx, const double *amp_tab, const int *step,
                             const double *sx2, const double *sy2)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  const double amp = amp_tab[*step];
  const long p = L.org + i * L.st[0] + j * L.st[1] + k * L.st[2];
  const double c = deform_corner(i, j, sx2, sy2, amp);
  g0[p] = R(c - deform_corner(i, j - 1, sx2, sy2, amp));
  g1[p] = R(deform_corner(i - 1, j, sx2, sy2, amp) - c);
  if (L.nd == 3) g2[p] = R(0);
}

// pairs of planes along one dimension compared by k_pairs_differ
struct PlanePairs {
  int n;
  int pa[5], pb[5];
};

// flag[0] = tag * 16 + q + 1 where a's planes pa[q] and pb[q] along
// dimension d differ bit for bit, for a row (i, j) of [i0, i0 + ni) x
// [j0, j0 + nj) over the other two dimensions in order (any one of the
// differing rows); flag[1], flag[2] = its i and j
template <class R>
__global__ void k_pairs_differ(const R *a, Layout L, int d, int i0, int ni, int j0, int nj, PlanePairs pp,
                               int tag, int *flag)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long rows = (long)ni * nj;
  if (t >= rows * pp.n) return;
  const int q = (int)(t / rows);
  const long r = t - q * rows;
  const int i = i0 + (int)(r / nj), j = j0 + (int)(r % nj);
  const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
  const long p = L.org + (long)i * L.st[d1] + (long)j * L.st[d2];
  using B = typename std::conditional<sizeof(R) == 4, unsigned, unsigned long long>::type;
  if (__builtin_bit_cast(B, a[p + (long)pp.pa[q] * L.st[d]]) != __builtin_bit_cast(B, a[p + (long)pp.pb[q] * L.st[d]])) {
    flag[0] = tag * 16 + q + 1;
    flag[1] = i;
    flag[2] = j;
  }
}

// explicit forcing (solvers/mpdata_rhs.hpp:52-93): zero the forcing array
// over a box, and psi += dt_arg * R over a box
template <class R>
__global__ void k_set_box(R *a, Layout L, Box bx, R v)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  a[L.org + i * L.st[0] + j * L.st[1] + k * L.st[2]] = v;
}

template <class R>
__global__ void k_axpy_box(R *y, const R *x, Layout L, Box bx, R alpha)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  const long p = L.org + i * L.st[0] + j * L.st[1] + k * L.st[2];
  y[p] += alpha * x[p];
}

static __global__ void k_tick(int *step, int set)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *step = set >= 0 ? set : *step + 1;
}

template <class R>
__global__ void k_fill_value(R *a, long n, R v)
{
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// var_dt output: the record interpolated in time (output_common.hpp:
// 112-137), element by element over whole arrays, each product and sum
// rounded on its own like the reference's array statements.  Linear:
// y = y * c0 + c1 * psi with c0 = (t1 - t) / (t1 - t0), c1 = (t - t0) /
// (t1 - t0).  Quadratic through (t0, y0), (t1, y1), (t2, y2) with the time
// differences d10 = t1 - t0, d20 = t2 - t0, d21 = t2 - t1, e0 = t - t0,
// e1 = t - t1 formed in double on the host.
template <class R>
__global__ void k_out_lin(R *y, const R *psi, long n, R c0, R c1)
{
#pragma clang fp contract(off)
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const R a = y[q] * c0;
  y[q] = a + c1 * psi[q];
}

template <class R>
__global__ void k_out_quad(R *y0, const R *y1, const R *y2, long n, R d10, R d20, R d21, R e0, R e1)
{
#pragma clang fp contract(off)
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const R a = y0[q], b = y1[q], c = y2[q];
  y0[q] = a + (b - a) / d10 * e0 + ((c - b) / (d21 * d20) - (b - a) / (d10 * d20)) * e0 * e1;
}

// ---------------------------------------------------------------- copies and reductions

template <class R>
__global__ void k_copy_box(R *dst, long dorg, long ds0, long ds1, long ds2,
                           const R *src, long sorg, long ss0, long ss1, long ss2, Box bx)
{
  int i, j, k;
  if (!box_idx(bx, i, j, k)) return;
  dst[dorg + i * ds0 + j * ds1 + k * ds2] = src[sorg + i * ss0 + j * ss1 + k * ss2];
}

}  // namespace mpdx

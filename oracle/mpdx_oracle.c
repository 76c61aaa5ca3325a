/*
 * mpdx_oracle.c -- TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).
 *
 * A loop-level FP64 restatement of one libmpdata++ MPDATA(-FCT) time step as
 * run by concurr::serial, kept in the reference's own unfused per-pass
 * structure: every halo fill is replayed with the reference's ranges and in
 * the reference's order, every formula keeps the reference's
 * parenthesisation.  All arrays start NaN-filled (as the reference's debug
 * build does, libmpdata++/blitz.hpp:107-112) so that a read of a value the
 * reference never defines poisons the result instead of passing silently.
 *
 * Citations are path:line in the reference tree (libmpdata++/...).
 *
 * Pinning: the reference's own output files pin donor cell, the standard and
 * iga antidiffusive velocities, FCT, open halos and the forcing schemes;
 * exact identities and textbook reductions pin cyclic halos, n_iters >= 3,
 * abs, npa, pfc, khn, hint_scale, upwind_filter_freq, var_dt and its
 * interpolated output (exact on fields linear / quadratic in time); an
 * independent evaluation of its closed form pins the deformational advector
 * generator.
 *
 * The heavy sweeps carry OpenMP pragmas so the same code doubles as the
 * multi-core CPU baseline; each element is computed independently, so the
 * result does not depend on the thread count.
 */
#include "mpdx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* libmpdata++/opts.hpp:34-49 */
enum {
  O_FCT = 1u << 0, O_ABS = 1u << 1, O_TOT = 1u << 2, O_PFC = 1u << 3,
  O_NPA = 1u << 4, O_IGA = 1u << 5, O_NUG = 1u << 6, O_DFL = 1u << 7,
  O_KHN = 1u << 8, O_DIV2 = 1u << 9, O_DIV3 = 1u << 10, O_DIV3DT = 1u << 11,
  O_FOT = 1u << 12
};
enum { BC_CYCLIC = 0, BC_OPEN = 1 };

typedef struct { int f, l; } rng;

struct orc {
  orc_params_t p;
  int nd, hw;
  int n[3];
  long sh[3], st[3], org, size;
  double *psi[ORC_MAX_EQNS][2];
  int lev[ORC_MAX_EQNS];
  double *GC[3], *tmp[2][3], *flux[3], *mono[3], *khn[3];
  double *psi_min, *psi_max, *beta_up, *beta_dn;
  long timestep;
  double time;
  int panic;
  int rhs_scheme;             /* 0 none, 1 euler_a, 2 euler_b, 3 trapez */
  orc_rhs_fn rhs_fn;
  void *rhs_user;
  double *rhs[ORC_MAX_EQNS];
  int gen_kind;               /* 0 none, 1 reversing deformational flow */
  double gen_T, gen_scale;
  double *gen_sx2, *gen_sy2;  /* sin^2 at corners a + 1/2 from a = -2 */
  /* var_dt output: intrp_vars (slot k of output variable v), do_record_cnt,
     intrp_times, record_time (output_common.hpp:32-37) */
  orc_out_fn out_fn;
  void *out_user;
  double out_freq;
  int out_ord, out_n, out_var[ORC_MAX_EQNS];
  double *out_buf[2][ORC_MAX_EQNS];
  int out_cnt;
  double out_t[2], out_rec;
};

static inline rng R_(int f, int l) { rng r = {f, l}; return r; }
static inline rng ext_(rng r, int n) { return R_(r.f - n, r.l + n); } /* r^n  (arakawa_c.hpp:43-51) */
static inline rng h_(rng r) { return R_(r.f - 1, r.l); }               /* r^h */

static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* pospart/negpart, formulae/common.hpp:87-124 */
static inline double pospart(unsigned o, double x) { return (o & O_NPA) ? (x + fabs(x)) / 2 : dmax(0, x); }
static inline double negpart(unsigned o, double x) { return (o & O_NPA) ? (x - fabs(x)) / 2 : dmin(0, x); }

/* frac / fct_frac, formulae/mpdata/formulae_mpdata_common.hpp:61-98 */
static inline double frac(unsigned o, double nom, double den)
{
  if (o & O_PFC) return den != 0 ? nom / den : 0;
  return nom / (den + DBL_MIN);
}
static inline double fct_frac(double nom, double den) { return nom / (den + DBL_EPSILON); }

/* donor-cell flux F, formulae/donorcell_formulae.hpp:25-34 */
static inline double F(unsigned o, double psi_l, double psi_r, double C)
{
  return pospart(o, C) * psi_l + negpart(o, C) * psi_r;
}

/* ---------------------------------------------------------------- indexing */

static inline long ix(const orc_t *s, int i, int j, int k)
{
  return s->org + (long)i * s->st[0] + (long)j * s->st[1] + (long)k * s->st[2];
}

/* rotated index: position[d]=a, position[d+1]=b, position[d+2]=c
 * (idxperm::pi<d>, formulae/idxperm.hpp:96-125) */
typedef struct { long o, sA, sB, sC; } rot;
static inline rot rot_(const orc_t *s, int d)
{
  rot r;
  r.o = s->org;
  r.sA = s->st[d];
  r.sB = s->st[(d + 1) % s->nd];
  r.sC = s->nd == 3 ? s->st[(d + 2) % 3] : 0;
  return r;
}
#define RR(r, a, b, c) ((r).o + (long)(a) * (r).sA + (long)(b) * (r).sB + (long)(c) * (r).sC)

static rng interior(const orc_t *s, int d) { return R_(0, s->n[d] - 1); }

/* ---------------------------------------------------------------- halos */

/* scalar halo fill for the bcond on dimension d over the rotated ranges
 * (ra, rb): cyclic copies with period N-1 (bcond/cyclic_3d.hpp:29-33,120-124),
 * open copies the edge value (bcond/open_3d.hpp:32-42,154-164). Ranges of
 * bcond_common.hpp:233-254. */
static void fill_sclr(orc_t *s, double *a, int d, int side, rng ra, rng rb)
{
  const int N = s->n[d], hw = s->hw, kind = s->p.bc[d][side];
  rot r = rot_(s, d);
  if (s->nd == 2) rb = R_(0, 0);
  for (int b = ra.f; b <= ra.l; ++b)
    for (int c = rb.f; c <= rb.l; ++c)
      for (int m = 1; m <= hw; ++m) {
        if (side == 0)
          a[RR(r, -m, b, c)] = a[RR(r, kind == BC_CYCLIC ? N - 1 - m : 0, b, c)];
        else
          a[RR(r, N - 1 + m, b, c)] = a[RR(r, kind == BC_CYCLIC ? m : N - 1, b, c)];
      }
}

/* tangential advector halo: cyclic = scalar copy (cyclic_3d.hpp:73-76),
 * open = zero over the scalar halo (open_3d.hpp:127-133,248-254) */
static void fill_nrml(orc_t *s, double *a, int d, int side, rng ra, rng rb)
{
  if (s->p.bc[d][side] == BC_CYCLIC) { fill_sclr(s, a, d, side, ra, rb); return; }
  const int N = s->n[d], hw = s->hw;
  rot r = rot_(s, d);
  if (s->nd == 2) rb = R_(0, 0);
  for (int b = ra.f; b <= ra.l; ++b)
    for (int c = rb.f; c <= rb.l; ++c)
      for (int m = 1; m <= hw; ++m)
        a[RR(r, side == 0 ? -m : N - 1 + m, b, c)] = 0;
}

/* normal advector halo (fill_halos_vctr_alng).
 * cyclic: cyclic_3d.hpp:44-48,135-139 (period N-1, `ad` ignored).
 * open:   open_3d.hpp:73-125,196-246 / open_2d.hpp fill_halos_vctr_alng:
 *         zero some perpendicular components at the edge, then set the halo
 *         faces so the edge cell is divergence free; `ad` skips the innermost
 *         halo face. Serial run: thread_rank == 0 == thread_size-1. */
static void fill_alng(orc_t *s, double **av, int d, int side, rng ra, rng rb, int ad)
{
  const int N = s->n[d], hw = s->hw, nd = s->nd;
  const int d1 = (d + 1) % nd, d2 = nd == 3 ? (d + 2) % 3 : -1;
  rot r = rot_(s, d);
  if (nd == 2) rb = R_(0, 0);
  if (s->p.bc[d][side] == BC_CYCLIC) {
    for (int b = ra.f; b <= ra.l; ++b)
      for (int c = rb.f; c <= rb.l; ++c)
        for (int m = 1; m <= hw; ++m) {
          if (side == 0) av[d][RR(r, -m, b, c)] = av[d][RR(r, N - 1 - m, b, c)];
          else           av[d][RR(r, N - 2 + m, b, c)] = av[d][RR(r, m - 1, b, c)];
        }
    return;
  }
  const int e = side == 0 ? 0 : N - 1;
  if (nd == 3) {
    if (d == 0) {
      for (int b = ra.f; b <= ra.l; ++b) {
        av[d2][RR(r, e, b, rb.f - 1)] = 0;
        av[d2][RR(r, e, b, rb.l)] = 0;
      }
      for (int c = rb.f; c <= rb.l; ++c) {
        av[d1][RR(r, e, ra.f - 1, c)] = 0;
        av[d1][RR(r, e, ra.l, c)] = 0;
      }
    } else if (d == 1) {
      for (int c = rb.f; c <= rb.l; ++c) {
        av[d1][RR(r, e, ra.f - 1, c)] = 0;
        av[d1][RR(r, e, ra.l, c)] = 0;
      }
    }
  } else if (d == 0) {
    av[d1][RR(r, e, ra.f - 1, 0)] = 0;
    av[d1][RR(r, e, ra.l, 0)] = 0;
  }
  const int lo = side == 0 ? -hw : N - 1 + (ad ? 1 : 0);
  const int hi = side == 0 ? -1 - (ad ? 1 : 0) : N - 2 + hw;
  for (int ii = lo; ii <= hi; ++ii)
    for (int b = ra.f; b <= ra.l; ++b)
      for (int c = rb.f; c <= rb.l; ++c) {
        if (nd == 3) {
          if (side == 0)
            av[d][RR(r, ii, b, c)] = av[d][RR(r, e, b, c)]
              - (av[d1][RR(r, e, b - 1, c)] - av[d1][RR(r, e, b, c)])
              - (av[d2][RR(r, e, b, c - 1)] - av[d2][RR(r, e, b, c)]);
          else
            av[d][RR(r, ii, b, c)] = av[d][RR(r, e - 1, b, c)]
              + (av[d1][RR(r, e, b - 1, c)] - av[d1][RR(r, e, b, c)])
              + (av[d2][RR(r, e, b, c - 1)] - av[d2][RR(r, e, b, c)]);
        } else {
          if (side == 0)
            av[d][RR(r, ii, b, 0)] = av[d][RR(r, e, b, 0)]
              - (av[d1][RR(r, e, b - 1, 0)] - av[d1][RR(r, e, b, 0)]);
          else
            av[d][RR(r, ii, b, 0)] = av[d][RR(r, e - 1, b, 0)]
              + (av[d1][RR(r, e, b - 1, 0)] - av[d1][RR(r, e, b, 0)]);
        }
      }
}

/* xchng_sclr: solver_3d.hpp:47-67 (y, then x, then z), solver_2d.hpp:40-56 */
static void xchng_sclr(orc_t *s, double *a, int ext)
{
  rng X = ext_(interior(s, 0), ext), Y = ext_(interior(s, 1), ext);
  if (s->nd == 3) {
    rng Z = ext_(interior(s, 2), ext);
    for (int sd = 0; sd < 2; ++sd) fill_sclr(s, a, 1, sd, Z, X);
    for (int sd = 0; sd < 2; ++sd) fill_sclr(s, a, 0, sd, Y, Z);
    for (int sd = 0; sd < 2; ++sd) fill_sclr(s, a, 2, sd, X, Y);
  } else {
    for (int sd = 0; sd < 2; ++sd) fill_sclr(s, a, 0, sd, Y, Y);
    for (int sd = 0; sd < 2; ++sd) fill_sclr(s, a, 1, sd, X, X);
  }
}

/* xchng_vctr_alng: solver_3d.hpp:73-103, solver_2d.hpp:63-88 */
static void xchng_vctr_alng(orc_t *s, double **av, int ad)
{
  rng I = interior(s, 0), J = interior(s, 1);
  if (s->nd == 3) {
    rng K = interior(s, 2);
    for (int sd = 0; sd < 2; ++sd) fill_alng(s, av, 0, sd, J, K, ad);
    for (int sd = 0; sd < 2; ++sd) fill_alng(s, av, 1, sd, K, I, ad);
    for (int sd = 0; sd < 2; ++sd) fill_alng(s, av, 2, sd, I, J, ad);
  } else {
    for (int sd = 0; sd < 2; ++sd) fill_alng(s, av, 0, sd, J, J, ad);
    for (int sd = 0; sd < 2; ++sd) fill_alng(s, av, 1, sd, I, I, ad);
  }
}

/* xchng_vctr_nrml: solver_3d.hpp:207-255, solver_2d.hpp:171-190 */
static void xchng_vctr_nrml(orc_t *s, double **av, int ext)
{
  rng I = interior(s, 0), J = interior(s, 1);
  if (s->nd == 3) {
    rng K = interior(s, 2);
    rng Xh = h_(ext_(I, ext)), X1 = ext_(ext_(I, ext), 1);
    rng Yh = h_(ext_(J, ext)), Y1 = ext_(ext_(J, ext), 1);
    rng Zh = h_(ext_(K, ext)), Z1 = ext_(ext_(K, ext), 1);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[0], 1, sd, Z1, Xh);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[0], 2, sd, Xh, Y1);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[1], 0, sd, Yh, Z1);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[1], 2, sd, X1, Yh);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[2], 0, sd, Y1, Zh);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[2], 1, sd, Zh, X1);
  } else {
    rng Xh = h_(ext_(I, ext)), Yh = h_(ext_(J, ext));
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[0], 1, sd, Xh, Xh);
    for (int sd = 0; sd < 2; ++sd) fill_nrml(s, av[1], 0, sd, Yh, Yh);
  }
}

/* ---------------------------------------------------------------- formulae */

/* make_flux over faces (-1-e .. N_d-1+e) x (interior^e)^2:
 * donorcell_formulae.hpp:50-79; used with e=0 in mpdata_osc_3d.hpp:157-163
 * and e=1 for the beta denominators in mpdata_fct_3d.hpp:155-160 */
static void make_flux(orc_t *s, const double *psi, double *const *C, double **out, int e)
{
  const unsigned o = s->p.opts;
  for (int d = 0; d < s->nd; ++d) {
    rot r = rot_(s, d);
    const int d1 = (d + 1) % s->nd, d2 = s->nd == 3 ? (d + 2) % 3 : -1;
    const int af = -1 - e, al = s->n[d] - 1 + e;
    const int bf = -e, bl = s->n[d1] - 1 + e;
    const int cf = s->nd == 3 ? -e : 0, cl = s->nd == 3 ? s->n[d2] - 1 + e : 0;
#pragma omp parallel for schedule(static)
    for (int a = af; a <= al; ++a)
      for (int b = bf; b <= bl; ++b)
        for (int c = cf; c <= cl; ++c) {
          const long p0 = RR(r, a, b, c);
          out[d][p0] = F(o, psi[p0], psi[RR(r, a + 1, b, c)], C[d][p0]);
        }
  }
}

/* donorcell_sum over the interior: donorcell_formulae.hpp:105-165 (3D
 * parenthesisation line 152, Kahan variant lines 154-164 via kahan_sum.hpp) */
static void donorcell(orc_t *s, double *psi_new, const double *psi_old, double *const *flx)
{
  const unsigned o = s->p.opts;
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  const long s0 = s->st[0], s1 = s->st[1], s2 = s->st[2];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) {
        const long p = ix(s, i, j, k);
        double f[6];
        f[0] = flx[0][p]; f[1] = flx[0][p - s0];
        f[2] = flx[1][p]; f[3] = flx[1][p - s1];
        if (nd == 3) { f[4] = flx[2][p]; f[5] = flx[2][p - s2]; }
        if (!(o & O_KHN)) {
          if (nd == 3)
            psi_new[p] = psi_old[p] + ((-f[0] + f[1]) + (-f[2] + f[3]) + (-f[4] + f[5])) / 1;
          else
            psi_new[p] = psi_old[p] + ((-f[0] + f[1]) + (-f[2] + f[3])) / 1;
        } else {
          /* kahan_zro then kahan_add of psi_old, -flx_1, flx_2, ... */
          double sum = 0, c = 0, y, t;
          const int nin = 1 + 2 * nd;
          double in[7];
          in[0] = psi_old[p];
          for (int q = 0; q < 2 * nd; ++q) in[1 + q] = (q % 2 == 0) ? -f[q] / 1 : f[q] / 1;
          for (int q = 0; q < nin; ++q) {
            y = in[q] - c;
            t = sum + y;
            c = (t - sum) - y;
            sum = t;
          }
          psi_new[p] = sum;
        }
      }
}

/* antidiffusive pseudo-velocity, standard form.
 * 3D: formulae_mpdata_3d.hpp:253-295 with ndx/ndy/ndz_psi (psi_3d.hpp:261-464),
 *     GC1_bar_xy / GC2_bar_xz (gc_3d.hpp:98-115,153-170), G_bar_x = 1.
 * 2D: formulae_mpdata_2d.hpp:241-278 with psi_2d.hpp:126-243, gc_2d.hpp:74-90.
 * Faces -1..N_d-1 along d, interior across (mpdata_osc_3d.hpp:85-128). */
static void antidiff(orc_t *s, double **V, const double *psi, double *const *U)
{
  const unsigned o = s->p.opts;
  const int iga = (o & O_IGA) != 0, ab = (o & O_ABS) != 0;
  for (int d = 0; d < s->nd; ++d) {
    rot r = rot_(s, d);
    const int d1 = (d + 1) % s->nd, d2 = s->nd == 3 ? (d + 2) % 3 : -1;
    const int al = s->n[d] - 1, bl = s->n[d1] - 1, cl = s->nd == 3 ? s->n[d2] - 1 : 0;
#define PS(a, b, c) (ab ? fabs(psi[RR(r, a, b, c)]) : psi[RR(r, a, b, c)])
#pragma omp parallel for schedule(static)
    for (int a = -1; a <= al; ++a)
      for (int b = 0; b <= bl; ++b)
        for (int c = 0; c <= cl; ++c) {
          const long p0 = RR(r, a, b, c);
          const double u = U[d][p0];
          double ndx, ndy;
          if (iga) {
            ndx = (s->nd == 3) ? psi[RR(r, a + 1, b, c)] - psi[p0]
                               : 2 * (psi[RR(r, a + 1, b, c)] - psi[p0]) / (1 + 1);
            ndy = (psi[RR(r, a + 1, b + 1, c)] + psi[RR(r, a, b + 1, c)]
                   - psi[RR(r, a + 1, b - 1, c)] - psi[RR(r, a, b - 1, c)]) / 4;
          } else {
            ndx = 2 * frac(o, PS(a + 1, b, c) - PS(a, b, c), PS(a + 1, b, c) + PS(a, b, c));
            ndy = frac(o,
                       PS(a + 1, b + 1, c) + PS(a, b + 1, c) - PS(a + 1, b - 1, c) - PS(a, b - 1, c),
                       PS(a + 1, b + 1, c) + PS(a, b + 1, c) + PS(a + 1, b - 1, c) + PS(a, b - 1, c));
          }
          const double g1 = (U[d1][RR(r, a + 1, b, c)] + U[d1][RR(r, a, b, c)]
                             + U[d1][RR(r, a + 1, b - 1, c)] + U[d1][RR(r, a, b - 1, c)]) / 4;
          if (s->nd == 3) {
            double ndz;
            if (iga)
              ndz = (psi[RR(r, a + 1, b, c + 1)] + psi[RR(r, a, b, c + 1)]
                     - psi[RR(r, a + 1, b, c - 1)] - psi[RR(r, a, b, c - 1)]) / 4;
            else
              ndz = frac(o,
                         PS(a + 1, b, c + 1) + PS(a, b, c + 1) - PS(a + 1, b, c - 1) - PS(a, b, c - 1),
                         PS(a + 1, b, c + 1) + PS(a, b, c + 1) + PS(a + 1, b, c - 1) + PS(a, b, c - 1));
            const double g2 = (U[d2][RR(r, a + 1, b, c)] + U[d2][RR(r, a, b, c)]
                               + U[d2][RR(r, a + 1, b, c - 1)] + U[d2][RR(r, a, b, c - 1)]) / 4;
            V[d][p0] = fabs(u) / 2 * (1 - fabs(u) / 1) * ndx
                     - u / 2 * (g1 * ndy + g2 * ndz) / 1;
          } else {
            V[d][p0] = fabs(u) / 2 * (1 - fabs(u) / 1) * ndx
                     - u * g1 / (2 * 1) * ndy;
          }
        }
#undef PS
  }
}

/* psi_min / psi_max over interior^1: mpdata_fct_3d.hpp:31-55, mpdata_fct_2d.hpp:31-46 */
static void fct_init(orc_t *s, const double *psi)
{
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  const long s0 = s->st[0], s1 = s->st[1], s2 = s->st[2];
  const int kf = nd == 3 ? -1 : 0, kl = nd == 3 ? n2 : 0;
#pragma omp parallel for schedule(static)
  for (int i = -1; i <= n0; ++i)
    for (int j = -1; j <= n1; ++j)
      for (int k = kf; k <= kl; ++k) {
        const long p = ix(s, i, j, k);
        if (nd == 3) {
          s->psi_min[p] = dmin(dmin(dmin(dmin(dmin(dmin(psi[p], psi[p + s0]), psi[p - s0]),
                                               psi[p + s1]), psi[p - s1]), psi[p + s2]), psi[p - s2]);
          s->psi_max[p] = dmax(dmax(dmax(dmax(dmax(dmax(psi[p], psi[p + s0]), psi[p - s0]),
                                               psi[p + s1]), psi[p - s1]), psi[p + s2]), psi[p - s2]);
        } else {
          s->psi_min[p] = dmin(dmin(dmin(dmin(psi[p + s1], psi[p - s0]), psi[p]), psi[p + s0]), psi[p - s1]);
          s->psi_max[p] = dmax(dmax(dmax(dmax(psi[p + s1], psi[p - s0]), psi[p]), psi[p + s0]), psi[p - s1]);
        }
      }
}

/* beta_up / beta_dn over interior^1:
 * 3D formulae_mpdata_fct_3d.hpp:21-145, 2D formulae_mpdata_fct_2d.hpp:21-120 */
static void betas(orc_t *s, const double *psi, double *const *flx)
{
  const unsigned o = s->p.opts;
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  const long s0 = s->st[0], s1 = s->st[1], s2 = s->st[2];
  const int kf = nd == 3 ? -1 : 0, kl = nd == 3 ? n2 : 0;
#pragma omp parallel for schedule(static)
  for (int i = -1; i <= n0; ++i)
    for (int j = -1; j <= n1; ++j)
      for (int k = kf; k <= kl; ++k) {
        const long p = ix(s, i, j, k);
        double up_nom, dn_nom, up_den, dn_den;
        if (nd == 3) {
          up_nom = (dmax(s->psi_max[p], dmax(psi[p], dmax(psi[p + s0], dmax(psi[p - s0],
                    dmax(psi[p + s1], dmax(psi[p - s1], dmax(psi[p + s2], psi[p - s2])))))))
                    - psi[p]) * 1;
          dn_nom = (psi[p] - dmin(s->psi_min[p], dmin(psi[p + s1], dmin(psi[p - s0],
                    dmin(psi[p], dmin(psi[p + s0], dmin(psi[p + s2], dmin(psi[p - s2], psi[p - s1]))))))))
                    * 1;
          up_den = (pospart(o, flx[0][p - s0]) - negpart(o, flx[0][p]))
                 + (pospart(o, flx[1][p - s1]) - negpart(o, flx[1][p]))
                 + (pospart(o, flx[2][p - s2]) - negpart(o, flx[2][p]));
          dn_den = (pospart(o, flx[0][p]) - negpart(o, flx[0][p - s0]))
                 + (pospart(o, flx[1][p]) - negpart(o, flx[1][p - s1]))
                 + (pospart(o, flx[2][p]) - negpart(o, flx[2][p - s2]));
        } else {
          up_nom = (dmax(s->psi_max[p], dmax(psi[p + s1], dmax(psi[p - s0], dmax(psi[p],
                    dmax(psi[p + s0], psi[p - s1]))))) - psi[p]) * 1;
          dn_nom = (psi[p] - dmin(s->psi_min[p], dmin(psi[p + s1], dmin(psi[p - s0], dmin(psi[p],
                    dmin(psi[p + s0], psi[p - s1])))))) * 1;
          up_den = (pospart(o, flx[0][p - s0]) - negpart(o, flx[0][p]))
                 + (pospart(o, flx[1][p - s1]) - negpart(o, flx[1][p]));
          dn_den = (pospart(o, flx[0][p]) - negpart(o, flx[0][p - s0]))
                 + (pospart(o, flx[1][p]) - negpart(o, flx[1][p - s1]));
        }
        s->beta_up[p] = fct_frac(up_nom, up_den);
        s->beta_dn[p] = fct_frac(dn_nom, dn_den);
      }
}

/* limited antidiffusive velocity: formulae_mpdata_fct_3d.hpp:147-247,
 * formulae_mpdata_fct_2d.hpp:122-214 */
static void gc_mono(orc_t *s, const double *psi, double *const *V)
{
  const unsigned o = s->p.opts;
  const int absv = (o & O_ABS) && !(o & O_IGA);
  const double *bu = s->beta_up, *bd = s->beta_dn;
  for (int d = 0; d < s->nd; ++d) {
    rot r = rot_(s, d);
    const int d1 = (d + 1) % s->nd, d2 = s->nd == 3 ? (d + 2) % 3 : -1;
    const int al = s->n[d] - 1, bl = s->n[d1] - 1, cl = s->nd == 3 ? s->n[d2] - 1 : 0;
#pragma omp parallel for schedule(static)
    for (int a = -1; a <= al; ++a)
      for (int b = 0; b <= bl; ++b)
        for (int c = 0; c <= cl; ++c) {
          const long p0 = RR(r, a, b, c), p1 = RR(r, a + 1, b, c);
          const double v = V[d][p0];
          double m;
          const double m_dn_up = dmin(1, dmin(bd[p0], bu[p1]));
          const double m_up_dn = dmin(1, dmin(bu[p0], bd[p1]));
          if (!absv)
            m = v > 0 ? m_dn_up : m_up_dn;
          else if (v > 0)
            m = psi[p0] > 0 ? m_dn_up : m_up_dn;
          else
            m = psi[p1] > 0 ? m_up_dn : m_dn_up;
          s->mono[d][p0] = v * m;
        }
  }
}

/* max |div GC| over the interior: solver_3d.hpp:400-409, solver_2d.hpp:305-313 */
static double max_abs_div(orc_t *s)
{
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  const long s0 = s->st[0], s1 = s->st[1], s2 = s->st[2];
  double mx = -INFINITY;
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) {
        const long p = ix(s, i, j, k);
        double v;
        if (nd == 3)
          v = fabs((s->GC[0][p] - s->GC[0][p - s0]) + (s->GC[1][p] - s->GC[1][p - s1])
                   + (s->GC[2][p] - s->GC[2][p - s2])) / 1;
        else
          v = fabs((s->GC[0][p] - s->GC[0][p - s0]) + (s->GC[1][p] - s->GC[1][p - s1])) / 1;
        mx = dmax(mx, v);
      }
  return mx;
}

/* max Courant number over the interior, solver_3d.hpp:390-398 (2D:
 * solver_2d.hpp:295-303) */
static double courant_number(orc_t *s)
{
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  const long s0 = s->st[0], s1 = s->st[1], s2 = s->st[2];
  double mx = -INFINITY;
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) {
        const long p = ix(s, i, j, k);
        double v;
        if (nd == 3)
          v = 0.5 * (fabs(s->GC[0][p] + s->GC[0][p - s0]) + fabs(s->GC[1][p] + s->GC[1][p - s1])
                     + fabs(s->GC[2][p] + s->GC[2][p - s2])) / 1;
        else
          v = 0.5 * (fabs(s->GC[0][p] + s->GC[0][p - s0]) + fabs(s->GC[1][p] + s->GC[1][p - s1])) / 1;
        mx = dmax(mx, v);
      }
  return mx;
}

/* scale_gc, solver_3d.hpp:411-421 (solver_2d.hpp:315-324): the advector
 * view (faces -1/2 .. n-3/2 along d, interior across) times cur/old, then
 * its halos */
static void scale_gc(orc_t *s, double cur_dt, double old_dt)
{
  const int nd = s->nd;
  for (int d = 0; d < nd; ++d) {
    int lo[3] = {0, 0, 0}, hi[3] = {s->n[0] - 1, s->n[1] - 1, nd == 3 ? s->n[2] - 1 : 0};
    lo[d] = -1;
    hi[d] = s->n[d] - 2;
    for (int i = lo[0]; i <= hi[0]; ++i)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int k = lo[2]; k <= hi[2]; ++k) s->GC[d][ix(s, i, j, k)] *= cur_dt / old_dt;
  }
  xchng_vctr_alng(s, s->GC, 0);
  xchng_vctr_nrml(s, s->GC, s->hw - 1);
}

/* ---------------------------------------------------------------- MPDATA step */

/* GC_unco / GC_corr bookkeeping, solvers/detail/mpdata_common.hpp:40-60 */
static double **gc_unco(orc_t *s, int iter)
{
  return iter == 1 ? s->GC : (iter % 2 ? s->tmp[1] : s->tmp[0]);
}
static double **gc_corr(orc_t *s, int iter) { return iter % 2 ? s->tmp[0] : s->tmp[1]; }

/* fct_adjust_antidiff: mpdata_fct_3d.hpp:57-110, mpdata_fct_2d.hpp:48-91 */
static void fct_adjust(orc_t *s, const double *psi, int iter)
{
  double **V = gc_corr(s, iter);
  xchng_vctr_alng(s, V, 1);
  xchng_vctr_nrml(s, V, 0);
  double **flx;
  if (s->p.opts & O_IGA) flx = V;
  else { make_flux(s, psi, V, s->flux, 1); flx = s->flux; }
  betas(s, psi, flx);
  gc_mono(s, psi, V);
}

/* advop: mpdata_osc_3d.hpp:73-202, mpdata_osc_2d.hpp:70-183 */
static void advop(orc_t *s, int e)
{
  const unsigned o = s->p.opts;
  if (o & O_FCT) fct_init(s, s->psi[e][s->lev[e]]);
  for (int iter = 0; iter < s->p.n_iters; ++iter) {
    if (iter != 0) {
      s->lev[e] ^= 1;                                   /* cycle(e) */
      xchng_sclr(s, s->psi[e][s->lev[e]], s->hw);       /* xchng(e) */
      antidiff(s, gc_corr(s, iter), s->psi[e][s->lev[e]], gc_unco(s, iter));
      if (!(o & O_FCT) && iter != s->p.n_iters - 1)
        xchng_vctr_nrml(s, gc_corr(s, iter), 0);
      if (o & O_FCT) fct_adjust(s, s->psi[e][s->lev[e]], iter);
    }
    double **C = iter == 0 ? s->GC : ((o & O_FCT) ? s->mono : gc_corr(s, iter));
    double **flx;
    if (!(o & O_IGA) || iter == 0) { make_flux(s, s->psi[e][s->lev[e]], C, s->flux, 0); flx = s->flux; }
    else flx = C;
    donorcell(s, s->psi[e][s->lev[e] ^ 1], s->psi[e][s->lev[e]], flx);
    if (s->p.upwind_filter_freq > 0 && s->timestep % s->p.upwind_filter_freq == 0) break;
  }
}

/* scale by 2^-exp on the interior: solver_common.hpp:374-379 */
static void scale(orc_t *s, int e, int ex)
{
  if (ex == 0) return;
  const int nd = s->nd, n0 = s->n[0], n1 = s->n[1], n2 = nd == 3 ? s->n[2] : 1;
  double *a = s->psi[e][s->lev[e]];
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) {
        const long p = ix(s, i, j, k);
        if (ex > 0) a[p] /= (1 << ex);
        else a[p] *= (1 << -ex);
      }
}

static void seterr(char *err, int len, const char *msg)
{
  if (err && len > 0) { strncpy(err, msg, len - 1); err[len - 1] = 0; }
}

/* Time-varying advector of the reversing_deform test family
 * (tests/mp3_paper_2018_JCP/reversing_deform/reversing_deform.hpp:64-105):
 * the discrete curl of the streamfunction sin^2(pi x) sin^2(pi y) cos(pi t/T)
 * sampled at cell corners x = (a + 1/2)/(n0 - 1), y = (b + 1/2)/(n1 - 1)
 * (a and b taken modulo n - 1, the period of sin^2 in them),
 * each corner value rounded to a multiple of 2^-17 (so the discrete
 * divergence is exactly zero and the same values come out of any correctly
 * rounded evaluation), written over i, j = -1 .. n-1 at t + dt/2 and
 * followed by the advector halo refill of calc_gc_staggered (:98-100). */
static double deform_corner(const orc_t *s, int a, int b, double amp)
{
  return rint(amp * s->gen_sx2[a + 2] * s->gen_sy2[b + 2] * 131072.0) / 131072.0;
}

static void calc_gc_deform(orc_t *s)
{
  const double t_half = s->time + 0.5 * s->p.dt;
  const double amp = s->gen_scale * cos(3.141592653589793 * t_half / s->gen_T);
  const int k0 = 0, k1 = s->nd == 3 ? s->n[2] - 1 : 0;
  for (int i = -1; i <= s->n[0] - 1; ++i)
    for (int j = -1; j <= s->n[1] - 1; ++j)
      for (int k = k0; k <= k1; ++k) {
        const long p = ix(s, i, j, k);
        const double c = deform_corner(s, i, j, amp);
        s->GC[0][p] = c - deform_corner(s, i, j - 1, amp);
        s->GC[1][p] = deform_corner(s, i - 1, j, amp) - c;
        if (s->nd == 3) s->GC[2][p] = 0;
      }
  xchng_vctr_alng(s, s->GC, 0);
  xchng_vctr_nrml(s, s->GC, s->hw - 1);
}

int orc_set_generator(orc_t *s, int kind, double T, double max_courant)
{
  if (kind == 0) { s->gen_kind = 0; return 0; }
  if (s->p.var_dt) return 1;   /* its amplitude does not follow dt */
  if (kind != 1 || !(T > 0) || !(max_courant > 0)) return 1;
  const int n0 = s->n[0], n1 = s->n[1];
  const double pi = 3.141592653589793;
  free(s->gen_sx2);
  free(s->gen_sy2);
  s->gen_sx2 = (double *)malloc(sizeof(double) * (size_t)(n0 + 2));
  s->gen_sy2 = (double *)malloc(sizeof(double) * (size_t)(n1 + 2));
  /* each corner at its index modulo the period n-1 (the tables repeat bit
     for bit across the duplicated edge) */
  for (int a = -2; a < n0; ++a) {
    const int r = ((a % (n0 - 1)) + (n0 - 1)) % (n0 - 1);
    const double v = sin(pi * (r + 0.5) / (n0 - 1));
    s->gen_sx2[a + 2] = v * v;
  }
  for (int b = -2; b < n1; ++b) {
    const int r = ((b % (n1 - 1)) + (n1 - 1)) % (n1 - 1);
    const double v = sin(pi * (r + 0.5) / (n1 - 1));
    s->gen_sy2[b + 2] = v * v;
  }
  double mx = 0;
  for (int a = -1; a < n0; ++a)
    for (int b = -1; b < n1; ++b) {
      const double c = s->gen_sx2[a + 2] * s->gen_sy2[b + 2];
      mx = dmax(mx, fabs(c - s->gen_sx2[a + 2] * s->gen_sy2[b + 1]));
      mx = dmax(mx, fabs(s->gen_sx2[a + 1] * s->gen_sy2[b + 2] - c));
    }
  s->gen_kind = kind;
  s->gen_T = T;
  s->gen_scale = mx > 0 ? max_courant / mx : 0;
  return 0;
}

/* mpdata_rhs (solvers/mpdata_rhs.hpp:52-195): update_rhs zeroes the
 * forcing and lets the user add R; apply_rhs adds dt_arg * R to the state
 * over the interior */
static int update_rhs(orc_t *s, double dt, int at)
{
  const int n0 = s->n[0], n1 = s->n[1], n2 = s->nd == 3 ? s->n[2] : 1;
  for (int e = 0; e < s->p.n_eqns; ++e)
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n2; ++k) s->rhs[e][ix(s, i, j, k)] = 0;
  return s->rhs_fn ? s->rhs_fn(s, dt, at, s->rhs_user) : 0;
}

static void apply_rhs(orc_t *s, double dt_arg)
{
  const int n0 = s->n[0], n1 = s->n[1], n2 = s->nd == 3 ? s->n[2] : 1;
  for (int e = 0; e < s->p.n_eqns; ++e) {
    double *a = s->psi[e][s->lev[e]];
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n2; ++k) {
          const long p = ix(s, i, j, k);
          a[p] += dt_arg * s->rhs[e][p];
        }
  }
}

static double *alloc_nan(long n);

int orc_set_rhs_scheme(orc_t *s, int scheme, orc_rhs_fn fn, void *user)
{
  if (scheme < 0 || scheme > 3) return 1;
  for (int e = 0; e < s->p.n_eqns && scheme; ++e)
    if (!s->rhs[e]) s->rhs[e] = alloc_nan(s->size);
  s->rhs_scheme = scheme;
  s->rhs_fn = fn;
  s->rhs_user = user;
  return 0;
}

int orc_set_rhs(orc_t *s, int e, const double *dense)
{
  if (e < 0 || e >= s->p.n_eqns || !s->rhs[e]) return 1;
  const int n0 = s->n[0], n1 = s->n[1], n2 = s->nd == 3 ? s->n[2] : 1;
  long q = 0;
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k, ++q) s->rhs[e][ix(s, i, j, k)] = dense[q];
  return 0;
}

/* hook_ante_loop chain: solver_common.hpp:191-211 (advector halos along;
 * var_dt: rescale dt and the advector to max_courant), solver_3d.hpp:375-388
 * (divergence check), mpdata_osc_3d.hpp:40-50 (tangential advector halos,
 * ext = halo-1), mpdata_rhs.hpp:116-133 (trapez: R^0) */
static int ante_loop(orc_t *s, char *err, int errlen)
{
  xchng_vctr_alng(s, s->GC, 0);
  if (s->p.var_dt) {
    const double cfl = courant_number(s);
    if (cfl > 0) {
      const double prev_dt = s->p.dt;
      s->p.dt *= s->p.max_courant / cfl;
      scale_gc(s, s->p.dt, prev_dt);
    }
  }
  const double eps = s->p.max_abs_div_eps < 0 ? DBL_EPSILON : s->p.max_abs_div_eps;
  if (max_abs_div(s) > eps) {
    seterr(err, errlen, "libmpdata++: initial advector field is divergent");
    return 1;
  }
  xchng_vctr_nrml(s, s->GC, s->hw - 1);
  if (s->rhs_scheme == 3 && update_rhs(s, s->p.dt / 2, 0)) {
    seterr(err, errlen, "libmpdata++: update_rhs failed");
    return 1;
  }
  return 0;
}

static int solve(orc_t *s, long nt, double t_end, int by_time, char *err, int errlen);

/* solve(): solver_common.hpp:270-358; nt steps (with var_dt too, each with
   the current dt) */
int orc_advance(orc_t *s, long nt, char *err, int errlen)
{
  return solve(s, s->timestep + nt, 0, 0, err, errlen);
}

/* advance(real_t) of a var_dt solver: solver_common.hpp:292-355 steps while
   time < t_end or additional_steps > 0, which starts at out_intrp_ord (1 by
   default) and drops on each step that ends at or past t_end, so at least one
   step is taken */
int orc_advance_time(orc_t *s, double duration, char *err, int errlen)
{
  if (!s->p.var_dt) {
    seterr(err, errlen, "libmpdata++: advance by a duration needs var_dt");
    return 1;
  }
  return solve(s, 0, s->time + duration, 1, err, errlen);
}

double orc_dt(const orc_t *s) { return s->p.dt; }

/* ---- var_dt output: output_common.hpp:47-62,72-140 */

static int out_call(orc_t *s, int kind, double t, char *err, int errlen)
{
  if (s->out_fn(s, kind, t, s->out_user)) {
    seterr(err, errlen, "libmpdata++: output callback failed");
    return 1;
  }
  return 0;
}

/* intrp_vars[v + k n_eqns](ijk) = advectee(v)(ijk) */
static void out_save(orc_t *s, int k)
{
  for (int v = 0; v < s->out_n; ++v) {
    const int e = s->out_var[v];
    memcpy(s->out_buf[k][v], s->psi[e][s->lev[e]], sizeof(double) * (size_t)s->size);
  }
}

/* hook_ante_step, output_common.hpp:72-101 */
static void out_ante(orc_t *s)
{
  const double next_time = s->time + s->p.dt;
  const long next_idx = (long)floor(next_time / s->out_freq);
  const long curr_idx = (long)floor(s->time / s->out_freq);
  s->out_cnt = s->out_cnt > 0 ? s->out_cnt - 1 : 0;
  if (next_idx > curr_idx || s->out_cnt > 0) {
    if (s->out_cnt == 0) {
      s->out_cnt = s->out_ord;
      s->out_rec = next_idx * s->out_freq;
    }
    const int inc_ix = s->out_ord - s->out_cnt;
    s->out_t[inc_ix] = s->time;
    out_save(s, inc_ix);
  }
}

/* hook_post_step, output_common.hpp:104-156 (interpolation over the
   interior, then record_all) */
static int out_post(orc_t *s, char *err, int errlen)
{
  if (s->out_cnt != 1) return 0;
  const int n0 = s->n[0], n1 = s->n[1], n2 = s->nd == 3 ? s->n[2] : 1;
  const double t = s->out_rec;
  for (int v = 0; v < s->out_n; ++v) {
    const int e = s->out_var[v];
    double *y = s->out_buf[0][v];
    const double *psi = s->psi[e][s->lev[e]];
    for (int i = 0; i < n0; ++i)
      for (int j = 0; j < n1; ++j)
        for (int k = 0; k < n2; ++k) {
          const long q = ix(s, i, j, k);
          if (s->out_ord == 1) {
            const double t0 = s->out_t[0], t1 = s->time;
            y[q] *= (t1 - t) / (t1 - t0);
            y[q] += (t - t0) / (t1 - t0) * psi[q];
          } else {
            const double t0 = s->out_t[0], t1 = s->out_t[1], t2 = s->time;
            const double y0 = y[q], y1 = s->out_buf[1][v][q], y2 = psi[q];
            y[q] = y0 + (y1 - y0) / (t1 - t0) * (t - t0) +
                   ((y2 - y1) / ((t2 - t1) * (t2 - t0)) - (y1 - y0) / ((t1 - t0) * (t2 - t0))) * (t - t0) * (t - t1);
          }
        }
  }
  return out_call(s, 1, t, err, errlen);
}

int orc_set_var_dt_output(orc_t *s, double outfreq, int intrp_ord, int n_vars, const int *vars, orc_out_fn fn,
                          void *user)
{
  if (!fn) { s->out_fn = NULL; return 0; }
  if (!s->p.var_dt || !(outfreq > 0) || (intrp_ord != 1 && intrp_ord != 2) || n_vars < 1 ||
      n_vars > s->p.n_eqns)
    return 1;
  for (int v = 0; v < n_vars; ++v)
    if (vars[v] < 0 || vars[v] >= s->p.n_eqns) return 1;
  for (int k = 0; k < intrp_ord; ++k)
    for (int v = 0; v < n_vars; ++v)
      if (!s->out_buf[k][v]) s->out_buf[k][v] = alloc_nan(s->size);
  for (int v = 0; v < n_vars; ++v) s->out_var[v] = vars[v];
  s->out_n = n_vars;
  s->out_ord = intrp_ord;
  s->out_freq = outfreq;
  s->out_cnt = 0;
  s->out_fn = fn;
  s->out_user = user;
  return 0;
}

static int copy_interior(const orc_t *s, double *arr, double *dense, int to_arr);

int orc_get_output(const orc_t *s, int e, double *dense)
{
  for (int v = 0; v < s->out_n; ++v)
    if (s->out_fn && s->out_var[v] == e) return copy_interior(s, s->out_buf[0][v], dense, 0);
  return 1;
}

static int solve(orc_t *s, long nt, double t_end, int by_time, char *err, int errlen)
{
  const int var_dt = s->p.var_dt;
  int stepped = 0;
  if (s->timestep == 0) {
    if (ante_loop(s, err, errlen)) return 1;
    /* output_common::hook_ante_loop: intrp_vars = advectee, start + record */
    if (s->out_fn) {
      out_save(s, 0);
      s->out_cnt = 0;
      s->out_rec = s->time;
      if (out_call(s, 0, by_time ? t_end : s->time + (nt - s->timestep) * s->p.dt, err, errlen)) return 1;
    }
  }
  while (by_time ? (s->time < t_end || !stepped) : s->timestep < nt) {
    stepped = 1;
    if (s->panic) break;
    /* calc_gc() at the top of each step (solver_common.hpp:306-308) */
    if (s->gen_kind == 1) calc_gc_deform(s);
    /* mpdata_rhs::hook_ante_step (mpdata_rhs.hpp:135-160) */
    if (s->rhs_scheme == 1 || s->rhs_scheme == 2) {
      if (update_rhs(s, s->p.dt, 0)) { seterr(err, errlen, "libmpdata++: update_rhs failed"); return 1; }
      if (s->rhs_scheme == 2) apply_rhs(s, s->p.dt);
    } else if (s->rhs_scheme == 3) {
      apply_rhs(s, s->p.dt / 2);
    }
    /* output_common::hook_ante_step after its parents' */
    if (s->out_fn) out_ante(s);
    for (int e = 0; e < s->p.n_eqns; ++e) {
      scale(s, e, s->p.hint_scale[e]);
      xchng_sclr(s, s->psi[e][s->lev[e]], s->hw);
      advop(s, e);
      s->lev[e] ^= 1;
      scale(s, e, -s->p.hint_scale[e]);
    }
    s->timestep++;
    s->time = var_dt ? s->time + s->p.dt : s->timestep * s->p.dt;
    /* mpdata_rhs::hook_post_step (mpdata_rhs.hpp:162-183) */
    if (s->rhs_scheme == 1) {
      apply_rhs(s, s->p.dt);
    } else if (s->rhs_scheme == 3) {
      if (update_rhs(s, s->p.dt / 2, 1)) { seterr(err, errlen, "libmpdata++: update_rhs failed"); return 1; }
      apply_rhs(s, s->p.dt / 2);
    }
    if (s->out_fn && out_post(s, err, errlen)) return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------- setup */

static double *alloc_nan(long n)
{
  double *a = (double *)malloc(sizeof(double) * (size_t)n);
  if (a) for (long q = 0; q < n; ++q) a[q] = NAN;
  return a;
}

orc_t *orc_create(const orc_params_t *p, char *err, int errlen)
{
  const unsigned unsupported = O_TOT | O_NUG | O_DFL | O_DIV2 | O_DIV3 | O_DIV3DT | O_FOT;
  if (p->n_dims != 2 && p->n_dims != 3) { seterr(err, errlen, "libmpdata++: n_dims must be 2 or 3"); return NULL; }
  for (int d = 0; d < p->n_dims; ++d)
    if (p->n[d] < 1) { seterr(err, errlen, "libmpdata++: bogus grid size"); return NULL; }
  if (p->opts & unsupported) { seterr(err, errlen, "libmpdata++: option not supported by this oracle"); return NULL; }
  if ((p->opts & O_ABS) && (p->opts & O_IGA)) { seterr(err, errlen, "libmpdata++: abs & iga options are mutually exclusive"); return NULL; }
  if (p->n_iters < 1) { seterr(err, errlen, "libmpdata++: n_iters must be positive"); return NULL; }
  if ((p->opts & O_IGA) && p->n_iters > 2) { seterr(err, errlen, "libmpdata++: infinite gauge option uses just one corrective step"); return NULL; }
  if (p->n_eqns < 1 || p->n_eqns > ORC_MAX_EQNS) { seterr(err, errlen, "libmpdata++: bad n_eqns"); return NULL; }

  orc_t *s = (orc_t *)calloc(1, sizeof(orc_t));
  s->p = *p;
  s->nd = p->n_dims;
  s->hw = (p->opts & O_FCT) ? 2 : 1;
  for (int d = 0; d < 3; ++d) s->n[d] = d < s->nd ? p->n[d] : 1;
  for (int d = 0; d < s->nd; ++d)
    if (s->n[d] < s->hw + 2) { seterr(err, errlen, "libmpdata++: grid too small for the halo"); free(s); return NULL; }
  for (int d = 0; d < 3; ++d) s->sh[d] = d < s->nd ? s->n[d] + 2 * s->hw : 1;
  s->st[2] = 1;
  s->st[1] = s->sh[2];
  s->st[0] = s->sh[1] * s->sh[2];
  s->size = s->sh[0] * s->sh[1] * s->sh[2];
  s->org = (long)s->hw * s->st[0] + (long)s->hw * s->st[1] + (s->nd == 3 ? s->hw : 0);
  if (s->nd == 2) s->st[2] = 0;
  for (int e = 0; e < p->n_eqns; ++e) {
    s->psi[e][0] = alloc_nan(s->size);
    s->psi[e][1] = alloc_nan(s->size);
  }
  for (int d = 0; d < s->nd; ++d) {
    s->GC[d] = alloc_nan(s->size);
    s->tmp[0][d] = alloc_nan(s->size);
    s->tmp[1][d] = alloc_nan(s->size);
    s->flux[d] = alloc_nan(s->size);
    s->mono[d] = alloc_nan(s->size);
  }
  s->psi_min = alloc_nan(s->size);
  s->psi_max = alloc_nan(s->size);
  s->beta_up = alloc_nan(s->size);
  s->beta_dn = alloc_nan(s->size);
  return s;
}

void orc_destroy(orc_t *s)
{
  if (!s) return;
  for (int e = 0; e < s->p.n_eqns; ++e) { free(s->psi[e][0]); free(s->psi[e][1]); }
  for (int d = 0; d < s->nd; ++d) {
    free(s->GC[d]); free(s->tmp[0][d]); free(s->tmp[1][d]); free(s->flux[d]); free(s->mono[d]);
  }
  free(s->psi_min); free(s->psi_max); free(s->beta_up); free(s->beta_dn);
  free(s->gen_sx2); free(s->gen_sy2);
  for (int e = 0; e < s->p.n_eqns; ++e) free(s->rhs[e]);
  for (int k = 0; k < 2; ++k)
    for (int v = 0; v < ORC_MAX_EQNS; ++v) free(s->out_buf[k][v]);
  free(s);
}

static int copy_interior(const orc_t *s, double *arr, double *dense, int to_arr)
{
  const int n0 = s->n[0], n1 = s->n[1], n2 = s->nd == 3 ? s->n[2] : 1;
  long q = 0;
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k, ++q) {
        const long p = ix(s, i, j, k);
        if (to_arr) arr[p] = dense[q]; else dense[q] = arr[p];
      }
  return 0;
}

int orc_set_advectee(orc_t *s, int e, const double *dense)
{
  if (e < 0 || e >= s->p.n_eqns) return 1;
  return copy_interior(s, s->psi[e][s->lev[e]], (double *)dense, 1);
}

int orc_get_advectee(const orc_t *s, int e, double *dense)
{
  if (e < 0 || e >= s->p.n_eqns) return 1;
  return copy_interior(s, s->psi[e][s->lev[e]], dense, 0);
}

/* the view of sharedmem.hpp:431-456 (2D) and 543-577 (3D): lower bounds
 * (-1, 0, 0); component d covers faces 0..N_d-2 of grid^(-1)^h after
 * distmem_ext adds x=-1 */
static void adv_bounds(const orc_t *s, int d, int lo[3], int hi[3])
{
  for (int q = 0; q < 3; ++q) { lo[q] = 0; hi[q] = q < s->nd ? s->n[q] - 1 : 0; }
  lo[0] = -1;
  if (d == 0) hi[0] = s->n[0] - 2;
  else hi[d] = s->n[d] - 2;
}

int orc_advector_extents(const orc_t *s, int d, int ext[3])
{
  int lo[3], hi[3];
  if (d < 0 || d >= s->nd) return 1;
  adv_bounds(s, d, lo, hi);
  for (int q = 0; q < 3; ++q) ext[q] = hi[q] - lo[q] + 1;
  return 0;
}

static int copy_advector(const orc_t *s, int d, double *dense, int to_arr)
{
  int lo[3], hi[3];
  if (d < 0 || d >= s->nd) return 1;
  adv_bounds(s, d, lo, hi);
  long q = 0;
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k, ++q) {
        const long p = ix(s, i, j, k);
        if (to_arr) s->GC[d][p] = dense[q]; else dense[q] = s->GC[d][p];
      }
  return 0;
}

int orc_set_advector(orc_t *s, int d, const double *dense) { return copy_advector(s, d, (double *)dense, 1); }
int orc_get_advector(const orc_t *s, int d, double *dense) { return copy_advector(s, d, dense, 0); }

long orc_timestep(const orc_t *s) { return s->timestep; }
double orc_time(const orc_t *s) { return s->time; }
int *orc_panic_ptr(orc_t *s) { return &s->panic; }

int orc_full_shape(const orc_t *s, int *hw, int shape[3])
{
  *hw = s->hw;
  for (int d = 0; d < 3; ++d) shape[d] = (int)s->sh[d];
  return 0;
}

int orc_get_full(const orc_t *s, int what, int idx, double *out)
{
  const double *a = NULL;
  if (what == 0 && idx >= 0 && idx < s->p.n_eqns) a = s->psi[idx][s->lev[idx]];
  if (what == 1 && idx >= 0 && idx < s->nd) a = s->GC[idx];
  if (!a) return 1;
  memcpy(out, a, sizeof(double) * (size_t)s->size);
  return 0;
}

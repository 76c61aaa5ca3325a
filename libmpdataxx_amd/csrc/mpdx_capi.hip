// C ABI of the MI355X MPDATA backend (include/mpdx.h).
//
// Host side: validates the configuration like the reference's constructors,
// owns the device arrays, replays solver_common::solve (solver_common.hpp:
// 270-358) and mpdata_osc::advop (mpdata_osc_3d.hpp:73-202) as a sequence of
// kernel launches on one HIP stream, and exposes the advectee/advector views.
#include "mpdx.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mpdx_solver.hpp"
#include "mpdx_perpass.hpp"
#include "mpdx_fused.hpp"
#include "mpdx_reduce.hpp"
#include "mpdx_tasks.hpp"
#include "mpdx_interior_list.hpp"

// the FP32 default-guard interior kernels are instantiated in mpdx_interior.hip
namespace mpdx {
#define MPDX_EXTERN(R, G, F, I, S, Z, T, W)                                                    \
  extern template __global__ void k_fz_interior<R, G, F, I, S, Z, T, W>(InteriorArgs<R>);      \
  extern template __global__ void k_task_list<InteriorTask<R, G, F, I, S, Z, T, W>>(           \
      const InteriorTask<R, G, F, I, S, Z, T, W> *, const unsigned *, int);
MPDX_INTERIOR_F32_FAST(MPDX_EXTERN)
#undef MPDX_EXTERN
#define MPDX_EXTERN_MIX(R, G, T, W)                                                            \
  extern template __global__ void k_task_list<InteriorMixTask<R, G, T, W>>(                     \
      const InteriorMixTask<R, G, T, W> *, const unsigned *, int);
MPDX_INTERIOR_MIX_F32_FAST(MPDX_EXTERN_MIX)
#undef MPDX_EXTERN_MIX
}  // namespace mpdx

using namespace mpdx;

namespace {

struct rng { int f, l; };
inline rng R_(int f, int l) { return rng{f, l}; }
inline rng ext_(rng r, int n) { return R_(r.f - n, r.l + n); }
inline rng h_(rng r) { return R_(r.f - 1, r.l); }

thread_local std::string g_create_error;

}  // namespace

// Fixed scratch of mpdx_reduce_noalloc: the block partials of one reduction
// (a (hi, lo) pair each), 8192 x 16 bytes = 128 KiB per GPU for the life of
// the process.  Host code never uses this name as an address: each device's
// copy is looked up once (hipGetSymbolAddress, in mpdx_create with that
// device current) and kept in g_scratch.
static constexpr int kScratchParts = 8192;
static __device__ __align__(256) double2 scratch[kScratchParts];

namespace {

std::mutex g_scratch_mu;
std::map<int, double2 *> g_scratch;   // device ordinal -> its scratch

// the scratch of the current device, looked up on its first use
int scratch_for_current_device(int *dev_out, std::string &err)
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    err = std::string("libmpdata++: HIP error: ") + hipGetErrorString(e);
    return MPDX_EHIP;
  }
  *dev_out = dev;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch.count(dev)) return MPDX_OK;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    err = std::string("libmpdata++: cannot locate the reduction scratch: ") + hipGetErrorString(e);
    return MPDX_EHIP;
  }
  g_scratch[dev] = static_cast<double2 *>(p);
  return MPDX_OK;
}

double2 *scratch_of(int dev)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  return it == g_scratch.end() ? nullptr : it->second;
}

}  // namespace


#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      s->err = std::string("libmpdata++: HIP error: ") + hipGetErrorString(e_) \
               + " (" #expr ")";                                            \
      return MPDX_EHIP;                                                     \
    }                                                                       \
  } while (0)

namespace {

constexpr int kBlock = 256;

template <class K, class... A>
void launch(hipStream_t st, long total, K kern, A... args)
{
  if (total <= 0) return;
  const long nb = (total + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kBlock), 0, st, args...);
}

// the same, timed under `name` when profiling is on
template <class K, class... A>
void plaunch(mpdx_solver *s, const char *name, long total, K kern, A... args)
{
  if (total <= 0) return;
  ProfScope ps(s, name);
  launch(s->stream, total, kern, args...);
}

// grid of a Box kernel (box_idx): x covers the (b, c) plane, y the a extent
inline dim3 box_grid(const Box &bx)
{
  return dim3((bx.plane + kBlock - 1) / kBlock, (unsigned)bx.ext[0]);
}

template <class K, class... A>
void blaunch(hipStream_t st, const Box &bx, K kern, A... args)
{
  if (bx.total <= 0) return;
  hipLaunchKernelGGL(kern, box_grid(bx), dim3(kBlock), 0, st, args...);
}

template <class K, class... A>
void pblaunch(mpdx_solver *s, const char *name, const Box &bx, K kern, A... args)
{
  if (bx.total <= 0) return;
  ProfScope ps(s, name);
  blaunch(s->stream, bx, kern, args...);
}

template <class T>
void launch_task_list(hipStream_t st, const void *tasks, const unsigned *first, int n, unsigned blocks)
{
  hipLaunchKernelGGL(k_task_list<T>, dim3(blocks), dim3(T::kThreads / T::kBlockY, T::kBlockY), 0, st,
                     static_cast<const T *>(tasks), first, n);
}

// records task t, whose own launch has grid g, into the solver's recorder
template <class T>
void record(mpdx_solver *s, T t, dim3 g)
{
  t.nx = g.x;
  t.ny = g.y;
  t.nz = g.z;
  if (t.blocks() == 0) return;
  RecOp &op = s->rec->next();
  op.fn = &launch_task_list<T>;
  op.tsize = sizeof(T);
  op.blocks = t.blocks();
  op.bytes.resize(sizeof(T));
  std::memcpy(op.bytes.data(), &t, sizeof(T));
}

// a group of mutually independent recorded launches (no-op when launching)
struct RecGroup {
  mpdx_solver *s;
  explicit RecGroup(mpdx_solver *s_) : s(s_) { if (s->rec) s->rec->begin(); }
  ~RecGroup() { if (s->rec) s->rec->end(); }
};

rng interior(const mpdx_solver *s, int d) { return R_(0, s->L.n[d] - 1); }

Box interior_box(const Layout &L)
{
  return make_box(0, L.n[0] - 1, 0, L.n[1] - 1, 0, L.nd == 3 ? L.n[2] - 1 : 0);
}

// ----------------------------------------------------------------- per-pass stepper

template <class R, class G>
struct PerPass {
  mpdx_solver *s;
  const Layout &L;
  hipStream_t st;
  unsigned o;

  PerPass(mpdx_solver *s_) : s(s_), L(s_->L), st(s_->stream), o(s_->p.opts) {}

  R *A(void *p) const { return static_cast<R *>(p); }
  R **av_of(void **arr, R *out[3]) const
  {
    for (int d = 0; d < 3; ++d) out[d] = d < L.nd ? A(arr[d]) : A(arr[0]);
    return out;
  }

  // Rows (x, y) whose z halo of the velocity the interior kernel's
  // z-boundary tiles compute (open z) or store themselves (cyclic z): the
  // z-direction fills skip x in zskip_x x y in zskip_y.  zskip: the normal
  // component's fill; zskip_nrml: the tangential components' too.
  bool zskip = false, zskip_nrml = false;
  rng zskip_x{0, -1}, zskip_y{0, -1};

  // the rows of ra x rb (x, y) outside the skipped ones: x bands over all
  // of rb, then y bands between them
  std::vector<std::pair<rng, rng>> zrows(rng ra, rng rb) const
  {
    std::vector<std::pair<rng, rng>> v;
    const rng mx = R_(std::max(ra.f, zskip_x.f), std::min(ra.l, zskip_x.l));
    const rng my = R_(std::max(rb.f, zskip_y.f), std::min(rb.l, zskip_y.l));
    if (mx.f > mx.l || my.f > my.l) {
      v.emplace_back(ra, rb);
      return v;
    }
    if (ra.f < mx.f) v.emplace_back(R_(ra.f, mx.f - 1), rb);
    if (mx.l < ra.l) v.emplace_back(R_(mx.l + 1, ra.l), rb);
    if (rb.f < my.f) v.emplace_back(mx, R_(rb.f, my.f - 1));
    if (my.l < rb.l) v.emplace_back(mx, R_(my.l + 1, rb.l));
    return v;
  }

  // Scalar-style fills along d (mode 0 cyclic / 1 edge copy from the bcond
  // kind, or 2 = zero).  Fills queue up in a batch that runs as one launch;
  // callers flush wherever a later fill reads what an earlier one wrote.
  FillBatch<R> fb{};

  void fill(R *a, int d, int side, rng ra, rng rb, bool nrml)
  {
    if (zskip && zskip_nrml && nrml && L.nd == 3 && d == 2) {
      for (const auto &r : zrows(ra, rb)) fill_rows(a, d, side, r.first, r.second, nrml);
      return;
    }
    fill_rows(a, d, side, ra, rb, nrml);
  }

  void fill_rows(R *a, int d, int side, rng ra, rng rb, bool nrml)
  {
    const int kind = s->p.bc[d][side];
    int mode = kind == MPDX_BC_CYCLIC ? 0 : (nrml ? 2 : 1);
    // ra spans the rotation's b axis, rb its c axis; the box puts the
    // memory-contiguous one last (fastest thread index)
    Rot r = make_rot(L, d);
    Box bx;
    int ord = 0;
    if (L.nd == 2) {
      bx = make_box(1, L.hw, 0, 0, ra.f, ra.l);
      ord = 1;
    } else if (d == 1) {
      // rotation (y, z, x): swap the tangential axes so that c is z
      std::swap(r.sB, r.sC);
      bx = make_box(rb.f, rb.l, 1, L.hw, ra.f, ra.l);
    } else {
      bx = make_box(ra.f, ra.l, 1, L.hw, rb.f, rb.l);
    }
    if (bx.total <= 0) return;
    if (fb.n == 6) flush_fills();
    const int q = fb.n++;
    fb.a[q] = a; fb.r[q] = r; fb.bx[q] = bx;
    fb.side[q] = side; fb.mode[q] = mode; fb.N[q] = L.n[d]; fb.ord[q] = ord;
  }

  void flush_fills()
  {
    if (fb.n == 0) return;
    if (s->rec) {
      RecGroup grp(s);
      for (int q = 0; q < fb.n; ++q) {
        FillTask<R> t{};
        t.a = fb.a[q]; t.r = fb.r[q]; t.bx = fb.bx[q];
        t.side = fb.side[q]; t.mode = fb.mode[q]; t.N = fb.N[q]; t.ord = fb.ord[q];
        record(s, t, box_grid(fb.bx[q]));
      }
      fb.n = 0;
      return;
    }
    unsigned gx = 0, gy = 0;
    for (int q = 0; q < fb.n; ++q) {
      gx = std::max(gx, box_grid(fb.bx[q]).x);
      gy = std::max(gy, (unsigned)fb.bx[q].ext[0]);
    }
    {
      ProfScope ps(s, "pp_halo");
      hipLaunchKernelGGL(k_fill_batch<R>, dim3(gx, gy, fb.n), dim3(kBlock), 0, s->stream, fb);
    }
    fb.n = 0;
  }

  // both sides of the normal-component halo along d (open: edge zeroing of
  // the perpendicular components first, then the zero-divergence faces);
  // the two sides touch disjoint cells, so each step covers both at once
  void alng(R **av, int d, rng ra, rng rb, int ad)
  {
    const int N = L.n[d], hw = L.hw, nd = L.nd;
    Rot r = make_rot(L, d);
    if (nd == 2) rb = R_(0, 0);
    R *avd = av[d], *av1 = av[(d + 1) % nd], *av2 = nd == 3 ? av[(d + 2) % 3] : av[0];
    int cyc[2], eopen[2], nopen = 0;
    for (int side = 0; side < 2; ++side) {
      cyc[side] = s->p.bc[d][side] == MPDX_BC_CYCLIC;
      if (!cyc[side]) eopen[nopen++] = side == 0 ? 0 : N - 1;
    }
    ProfScope ps(s, "pp_halo");
    // (the perpendicular-component zeroing does nothing along z)
    if (nopen > 0 && !(nd == 3 && d == 2)) {
      const long nz = std::max(ra.l - ra.f + 1, rb.l - rb.f + 1);
      if (s->rec) {
        RecGroup grp(s);
        for (int q = 0; q < nopen; ++q) {
          AlngZeroTask<R> t{};
          t.av1 = av1; t.av2 = av2; t.r = r; t.nd = nd; t.d = d; t.e = eopen[q];
          t.raf = ra.f; t.ral = ra.l; t.rbf = rb.f; t.rbl = rb.l;
          record(s, t, dim3((unsigned)((nz + kBlock - 1) / kBlock)));
        }
      } else {
        hipLaunchKernelGGL(k_alng_zero<R>, dim3((unsigned)((nz + kBlock - 1) / kBlock), nopen), dim3(kBlock), 0,
                           s->stream, av1, av2, r, nd, d, eopen[0], eopen[nopen - 1], ra.f, ra.l, rb.f, rb.l);
      }
    }
    // the faces themselves: the thread index runs along the box's c axis,
    // which for the y faces is x under the rotation (y, z, x); swap the
    // tangential axes (and the components that vary along them) so that it
    // runs along z, the contiguous one
    int cfirst = 0;
    if (nd == 3 && d == 1) {
      std::swap(r.sB, r.sC);
      std::swap(ra, rb);
      std::swap(av1, av2);
      cfirst = 1;
    }
    Box bx[2];
    for (int side = 0; side < 2; ++side) {
      if (cyc[side]) {
        bx[side] = make_box(1, hw, ra.f, ra.l, rb.f, rb.l);
      } else {
        const int lo = side == 0 ? -hw : N - 1 + (ad ? 1 : 0);
        const int hi = side == 0 ? -1 - (ad ? 1 : 0) : N - 2 + hw;
        bx[side] = make_box(lo, hi, ra.f, ra.l, rb.f, rb.l);
      }
    }
    if (zskip && nd == 3 && d == 2) {
      std::vector<std::pair<Box, int>> v;
      for (int side = 0; side < 2; ++side)
        for (const auto &b : zrows(ra, rb))
          v.emplace_back(make_box(bx[side].lo[0], bx[side].lo[0] + bx[side].ext[0] - 1, b.first.f, b.first.l,
                                  b.second.f, b.second.l), side);
      if (s->rec) {
        RecGroup grp(s);
        for (const auto &bt : v) record_alng(avd, av1, av2, r, bt.first, nd, bt.second, cyc[bt.second], N, cfirst);
        return;
      }
      boxlist_launch(v, [&](const BoxList &bl) {
        hipLaunchKernelGGL(k_alng_list<R>, dim3(boxlist_blocks(bl)), dim3(kBlock), 0, s->stream, avd,
                           (const R *)av1, (const R *)av2, r, bl, nd, cyc[0], cyc[1], N, cfirst);
      });
      return;
    }
    if (s->rec) {
      RecGroup grp(s);
      for (int side = 0; side < 2; ++side)
        if (bx[side].total > 0) record_alng(avd, av1, av2, r, bx[side], nd, side, cyc[side], N, cfirst);
      return;
    }
    const unsigned gx = std::max(box_grid(bx[0]).x, box_grid(bx[1]).x);
    const unsigned gy = (unsigned)std::max(bx[0].ext[0], bx[1].ext[0]);
    if (bx[0].total > 0 || bx[1].total > 0)
      hipLaunchKernelGGL(k_alng2<R>, dim3(gx, gy, 2), dim3(kBlock), 0, s->stream, avd, (const R *)av1,
                         (const R *)av2, r, bx[0], bx[1], nd, cyc[0], cyc[1], N, cfirst);
  }

  void record_alng(R *avd, R *av1, R *av2, const Rot &r, const Box &bx, int nd, int side, int cyc, int N,
                   int cfirst)
  {
    AlngTask<R> t{};
    t.avd = avd; t.av1 = av1; t.av2 = av2; t.r = r; t.bx = bx;
    t.nd = nd; t.side = side; t.cyc = cyc; t.N = N; t.cfirst = cfirst;
    record(s, t, box_grid(bx));
  }

  // solver_3d.hpp:47-67, solver_2d.hpp:40-56
  void xchng_sclr(R *a, int ext)
  {
    rng X = ext_(interior(s, 0), ext), Y = ext_(interior(s, 1), ext);
    if (L.nd == 3) {
      rng Z = ext_(interior(s, 2), ext);
      // each dimension's fill reads the halos the previous one wrote
      for (int sd = 0; sd < 2; ++sd) fill(a, 1, sd, Z, X, false);
      flush_fills();
      for (int sd = 0; sd < 2; ++sd) fill(a, 0, sd, Y, Z, false);
      flush_fills();
      for (int sd = 0; sd < 2; ++sd) fill(a, 2, sd, X, Y, false);
      flush_fills();
    } else {
      for (int sd = 0; sd < 2; ++sd) fill(a, 0, sd, Y, Y, false);
      flush_fills();
      for (int sd = 0; sd < 2; ++sd) fill(a, 1, sd, X, X, false);
      flush_fills();
    }
  }

  // solver_3d.hpp:73-103, solver_2d.hpp:63-88
  void xchng_vctr_alng(R **av, int ad)
  {
    rng I = interior(s, 0), J = interior(s, 1);
    if (L.nd == 3) {
      rng K = interior(s, 2);
      alng(av, 0, J, K, ad);
      alng(av, 1, K, I, ad);
      alng(av, 2, I, J, ad);
    } else {
      alng(av, 0, J, J, ad);
      alng(av, 1, I, I, ad);
    }
  }

  // solver_3d.hpp:207-255, solver_2d.hpp:171-190
  void xchng_vctr_nrml(R **av, int ext)
  {
    rng I = interior(s, 0), J = interior(s, 1);
    if (L.nd == 3) {
      rng K = interior(s, 2);
      rng Xh = h_(ext_(I, ext)), X1 = ext_(ext_(I, ext), 1);
      rng Yh = h_(ext_(J, ext)), Y1 = ext_(ext_(J, ext), 1);
      rng Zh = h_(ext_(K, ext)), Z1 = ext_(ext_(K, ext), 1);
      // a component's second fill reads the halo its first one wrote;
      // different components are independent, so the first fills of all
      // three run together, then the second ones
      for (int sd = 0; sd < 2; ++sd) fill(av[0], 1, sd, Z1, Xh, true);
      for (int sd = 0; sd < 2; ++sd) fill(av[1], 0, sd, Yh, Z1, true);
      for (int sd = 0; sd < 2; ++sd) fill(av[2], 0, sd, Y1, Zh, true);
      flush_fills();
      for (int sd = 0; sd < 2; ++sd) fill(av[0], 2, sd, Xh, Y1, true);
      for (int sd = 0; sd < 2; ++sd) fill(av[1], 2, sd, X1, Yh, true);
      for (int sd = 0; sd < 2; ++sd) fill(av[2], 1, sd, Zh, X1, true);
      flush_fills();
    } else {
      rng Xh = h_(ext_(I, ext)), Yh = h_(ext_(J, ext));
      for (int sd = 0; sd < 2; ++sd) fill(av[0], 1, sd, Xh, Xh, true);
      for (int sd = 0; sd < 2; ++sd) fill(av[1], 0, sd, Yh, Yh, true);
      flush_fills();
    }
  }

  void make_flux(R *psi, R **C, R **out, int e)
  {
    for (int d = 0; d < L.nd; ++d) {
      const int d1 = (d + 1) % L.nd, d2 = L.nd == 3 ? (d + 2) % 3 : 0;
      Box bx = make_box(-1 - e, L.n[d] - 1 + e, -e, L.n[d1] - 1 + e,
                        L.nd == 3 ? -e : 0, L.nd == 3 ? L.n[d2] - 1 + e : 0);
      pblaunch(s, "pp_flux", bx, k_flux<R>, out[d], (const R *)psi, (const R *)C[d], make_rot(L, d), bx, o);
    }
  }

  void antidiff(R **V, R *psi, R **U)
  {
    for (int d = 0; d < L.nd; ++d) {
      const int d1 = (d + 1) % L.nd, d2 = L.nd == 3 ? (d + 2) % 3 : 0;
      Box bx = make_box(-1, L.n[d] - 1, 0, L.n[d1] - 1, 0, L.nd == 3 ? L.n[d2] - 1 : 0);
      pblaunch(s, "pp_antidiff", bx, k_antidiff<R, G>, V[d], (const R *)psi, (const R *)U[d],
             (const R *)U[d1], (const R *)U[L.nd == 3 ? d2 : 0], make_rot(L, d), bx, L.nd, o);
    }
  }

  Box box_ext1() const
  {
    return make_box(-1, L.n[0], -1, L.n[1], L.nd == 3 ? -1 : 0, L.nd == 3 ? L.n[2] : 0);
  }

  // the GC_unco / GC_corr ping-pong arrays (of equation tmp_sel when the
  // equations run batched)
  void *(*vel())[3] { return s->tmp_sel > 0 ? s->tmp_eq[s->tmp_sel] : s->tmp; }
  R **unco(int iter, R *buf[3])
  {
    if (iter == 1) return av_of(s->GC, buf);
    return av_of(iter % 2 ? vel()[1] : vel()[0], buf);
  }
  R **corr(int iter, R *buf[3]) { return av_of(iter % 2 ? vel()[0] : vel()[1], buf); }

  // mpdata_fct_3d.hpp:57-110
  void fct_adjust(R *psi, int iter)
  {
    R *Vb[3], *Fb[3], *Mb[3];
    R **V = corr(iter, Vb);
    xchng_vctr_alng(V, 1);
    xchng_vctr_nrml(V, 0);
    R **flx = V;
    if (!(o & O_IGA)) { flx = av_of(s->flux, Fb); make_flux(psi, V, flx, 1); }
    Box b1 = box_ext1();
    pblaunch(s, "pp_betas", b1, k_betas<R, G>, A(s->bup), A(s->bdn), (const R *)psi, (const R *)A(s->pmin),
           (const R *)A(s->pmax), (const R *)flx[0], (const R *)flx[1],
           (const R *)flx[L.nd == 3 ? 2 : 0], L, b1, o);
    R **M = av_of(s->mono, Mb);
    for (int d = 0; d < L.nd; ++d) {
      const int d1 = (d + 1) % L.nd, d2 = L.nd == 3 ? (d + 2) % 3 : 0;
      Box bx = make_box(-1, L.n[d] - 1, 0, L.n[d1] - 1, 0, L.nd == 3 ? L.n[d2] - 1 : 0);
      pblaunch(s, "pp_gc_mono", bx, k_gc_mono<R>, M[d], (const R *)V[d], (const R *)psi,
             (const R *)A(s->bup), (const R *)A(s->bdn), make_rot(L, d), bx, o);
    }
  }

  // mpdata_osc_3d.hpp:73-202
  void advop(int e)
  {
    int &lev = s->lev[e];
    if (o & O_FCT) {
      Box b1 = box_ext1();
      pblaunch(s, "pp_fct_init", b1, k_fct_init<R>, A(s->pmin), A(s->pmax), (const R *)A(s->psi[e][lev]), L, b1);
    }
    for (int iter = 0; iter < s->p.n_iters; ++iter) {
      R *Cb[3], *Fb[3], *Ub[3], *Vb[3];
      if (iter != 0) {
        lev ^= 1;
        xchng_sclr(A(s->psi[e][lev]), L.hw);
        antidiff(corr(iter, Vb), A(s->psi[e][lev]), unco(iter, Ub));
        if (!(o & O_FCT) && iter != s->p.n_iters - 1) xchng_vctr_nrml(corr(iter, Vb), 0);
        if (o & O_FCT) fct_adjust(A(s->psi[e][lev]), iter);
      }
      R **C = iter == 0 ? av_of(s->GC, Cb) : ((o & O_FCT) ? av_of(s->mono, Cb) : corr(iter, Cb));
      R **flx = C;
      if (!(o & O_IGA) || iter == 0) { flx = av_of(s->flux, Fb); make_flux(A(s->psi[e][lev]), C, flx, 0); }
      Box bi = interior_box(L);
      pblaunch(s, "pp_donorcell", bi, k_donorcell<R>, A(s->psi[e][lev ^ 1]), (const R *)A(s->psi[e][lev]),
             (const R *)flx[0], (const R *)flx[1], (const R *)flx[L.nd == 3 ? 2 : 0], L, bi, o);
      if (s->p.upwind_filter_freq > 0 && s->timestep % s->p.upwind_filter_freq == 0) break;
    }
  }

  void scale(int e, int ex)
  {
    if (ex == 0) return;
    Box bi = interior_box(L);
    const R f = ex > 0 ? R(1) / R(1 << ex) : R(1 << -ex);
    pblaunch(s, "pp_scale", bi, k_scale<R>, A(s->psi[e][s->lev[e]]), L, bi, f);
  }

  void step()
  {
    for (int e = 0; e < s->p.n_eqns; ++e) {
      scale(e, s->p.hint_scale[e]);
      xchng_sclr(A(s->psi[e][s->lev[e]]), L.hw);
      advop(e);
      s->lev[e] ^= 1;
      scale(e, -s->p.hint_scale[e]);
    }
  }
};

// ----------------------------------------------------------------- helpers

// captured time steps hold kernel arguments (buffer addresses, launch
// shapes, the generator's kernel); whatever changes one of them drops both
void drop_step_graphs(mpdx_solver *s)
{
  for (int q = 0; q < mpdx_solver::kGraphs; ++q) {
    if (s->graph[q]) (void)hipGraphExecDestroy(s->graph[q]);
    s->graph[q] = nullptr;
    if (s->eqplan[q].dev) (void)hipFree(s->eqplan[q].dev);
    s->eqplan[q] = BatchPlan();
  }
}

bool panicked(const mpdx_solver *s)
{
  return s->panic || (s->panic_ext && *s->panic_ext);
}

int fill_nan(mpdx_solver *s, void *p)
{
  if (s->esz == 8)
    launch(s->stream, s->L.size, k_fill_value<double>, (double *)p, s->L.size, (double)NAN);
  else
    launch(s->stream, s->L.size, k_fill_value<float>, (float *)p, s->L.size, (float)NAN);
  return 0;
}

int dev_alloc(mpdx_solver *s, void **out, size_t bytes)
{
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    s->err = std::string("libmpdata++: device allocation failed: ") + hipGetErrorString(e);
    return MPDX_ENOMEM;
  }
  s->allocs.push_back(p);
  *out = p;
  return 0;
}

int alloc_array(mpdx_solver *s, void **out)
{
  int rc = dev_alloc(s, out, (size_t)s->L.size * s->esz);
  if (rc) return rc;
  return fill_nan(s, *out);
}

int alloc_missing(mpdx_solver *s, void **slot)
{
  return *slot ? 0 : alloc_array(s, slot);
}

int ensure_perpass(mpdx_solver *s)
{
  if (s->perpass_ready) return 0;
  int rc = 0;
  for (int d = 0; d < s->L.nd && !rc; ++d) {
    rc = rc || alloc_missing(s, &s->tmp[0][d]);
    rc = rc || alloc_missing(s, &s->tmp[1][d]);
    rc = rc || alloc_missing(s, &s->flux[d]);
    rc = rc || alloc_missing(s, &s->mono[d]);
  }
  rc = rc || alloc_missing(s, &s->pmin);
  rc = rc || alloc_missing(s, &s->pmax);
  rc = rc || alloc_missing(s, &s->bup);
  rc = rc || alloc_missing(s, &s->bdn);
  if (!rc) s->perpass_ready = true;
  return rc;
}

// fused path: two advector scratch sets (GC_corr ping-pong) and one scalar
// scratch per equation
int fused_alloc(mpdx_solver *s)
{
  if (!s->fz) s->fz = new mpdx_fused_state();
  int rc = 0;
  for (int d = 0; d < s->L.nd && !rc; ++d) {
    if (s->p.n_iters > 1) rc = rc || alloc_missing(s, &s->tmp[0][d]);
    if (s->p.n_iters > 2) rc = rc || alloc_missing(s, &s->tmp[1][d]);
  }
  for (int e = 0; e < s->p.n_eqns && !rc; ++e)
    if (s->p.n_iters > 1) rc = alloc_missing(s, &s->fz->psi_tmp[e]);
  return rc;
}

void fused_free(mpdx_solver *s)
{
  delete s->fz;
  s->fz = nullptr;
}

int ensure_staging(mpdx_solver *s, size_t bytes)
{
  if (s->staging_bytes >= bytes) return 0;
  if (s->staging) {
    (void)hipFree(s->staging);
    for (auto &p : s->allocs)
      if (p == s->staging) p = nullptr;
    s->staging = nullptr;
  }
  int rc = dev_alloc(s, &s->staging, bytes);
  if (!rc) s->staging_bytes = bytes;
  return rc;
}

// advector(d) view bounds: sharedmem.hpp:431-456 (2D), 543-577 (3D)
void adv_bounds(const mpdx_solver *s, int d, int lo[3], int hi[3])
{
  for (int q = 0; q < 3; ++q) { lo[q] = 0; hi[q] = q < s->L.nd ? s->L.n[q] - 1 : 0; }
  lo[0] = -1;
  if (d == 0) hi[0] = s->L.n[0] - 2;
  else hi[d] = s->L.n[d] - 2;
}

template <class R>
void launch_copy(hipStream_t st, R *dst, long dorg, const long ds[3], const R *src, long sorg,
                 const long ss[3], Box bx)
{
  blaunch(st, bx, k_copy_box<R>, dst, dorg, ds[0], ds[1], ds[2], src, sorg, ss[0], ss[1], ss[2], bx);
}

// copy between a user buffer and the box [lo, hi] of a padded device array
int copy_view(mpdx_solver *s, void *arr, const int lo[3], const int hi[3], void *buf,
              const int64_t strides[3], int on_device, bool to_dev)
{
  const int nd = s->L.nd;
  int ext[3];
  for (int q = 0; q < 3; ++q) ext[q] = hi[q] - lo[q] + 1;
  Box bx = make_box(0, ext[0] - 1, 0, ext[1] - 1, 0, ext[2] - 1);
  const long aorg = s->L.org + lo[0] * s->L.st[0] + lo[1] * s->L.st[1] + (nd == 3 ? lo[2] * s->L.st[2] : 0);
  long ust[3] = {(long)strides[0], (long)strides[1], nd == 3 ? (long)strides[2] : 0};
  const long cst[3] = {(long)ext[1] * ext[2], (long)ext[2], 1};
  const size_t n = (size_t)bx.total, esz = (size_t)s->esz;
  void *dbuf = buf;
  const long *bst = ust;
  std::vector<char> host;
  // a device buffer may come from work on any stream (the library's own
  // stream is non-blocking): wait for everything enqueued on the device
  if (on_device) HIPCHK(hipDeviceSynchronize());
  if (!on_device) {
    if (int rc = ensure_staging(s, n * esz)) return rc;
    host.resize(n * esz);
    if (to_dev) {
      const char *src = static_cast<const char *>(buf);
      size_t q = 0;
      for (int i = 0; i < ext[0]; ++i)
        for (int j = 0; j < ext[1]; ++j)
          for (int k = 0; k < ext[2]; ++k, ++q)
            std::memcpy(&host[q * esz], src + (i * ust[0] + j * ust[1] + k * ust[2]) * (long)esz, esz);
      HIPCHK(hipMemcpyAsync(s->staging, host.data(), n * esz, hipMemcpyHostToDevice, s->stream));
    }
    dbuf = s->staging;
    bst = cst;
  }
  if (s->esz == 8) {
    if (to_dev) launch_copy<double>(s->stream, (double *)arr, aorg, s->L.st, (const double *)dbuf, 0, bst, bx);
    else launch_copy<double>(s->stream, (double *)dbuf, 0, bst, (const double *)arr, aorg, s->L.st, bx);
  } else {
    if (to_dev) launch_copy<float>(s->stream, (float *)arr, aorg, s->L.st, (const float *)dbuf, 0, bst, bx);
    else launch_copy<float>(s->stream, (float *)dbuf, 0, bst, (const float *)arr, aorg, s->L.st, bx);
  }
  HIPCHK(hipGetLastError());
  if (!on_device && !to_dev) {
    HIPCHK(hipMemcpyAsync(host.data(), s->staging, n * esz, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    char *dst = static_cast<char *>(buf);
    size_t q = 0;
    for (int i = 0; i < ext[0]; ++i)
      for (int j = 0; j < ext[1]; ++j)
        for (int k = 0; k < ext[2]; ++k, ++q)
          std::memcpy(dst + (i * ust[0] + j * ust[1] + k * ust[2]) * (long)esz, &host[q * esz], esz);
  } else {
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return 0;
}

// reduction of kind K over the interior into out_dev (device memory),
// block partials in part (red_parts(n0) entries)
template <int K, class R>
void red_launch(hipStream_t st, double2 *part, double *out_dev, const R *a, const R *g1, const R *g2,
                const Layout &L)
{
  const Box bi = interior_box(L);
  const int S = red_split(bi.ext[0]);
  const int nparts = bi.ext[0] * S;
  hipLaunchKernelGGL((k_red_part<K, R>), dim3((unsigned)nparts), dim3(kRedThreads), 0, st, part, a, g1, g2, L, bi, S);
  hipLaunchKernelGGL((k_red_final<K>), dim3(1), dim3(kRedFinalThreads), 0, st, out_dev, (const double2 *)part,
                     nparts);
}

// min / max / FP64 compensated sum (op = MPDX_RED_*) of arr into out_dev
template <class R>
void red_field(hipStream_t st, double2 *part, double *out_dev, const R *arr, int op, const Layout &L)
{
  if (op == MPDX_RED_MIN) red_launch<RK_MIN, R>(st, part, out_dev, arr, arr, arr, L);
  else if (op == MPDX_RED_MAX) red_launch<RK_MAX, R>(st, part, out_dev, arr, arr, arr, L);
  else red_launch<RK_SUM, R>(st, part, out_dev, arr, arr, arr, L);
}

// the result of a reduction enqueued into s->red_out, on the host
int red_result(mpdx_solver *s, double *out)
{
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->red_out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

template <class R>
int reduce_impl(mpdx_solver *s, const void *arr, int op, double *out)
{
  {
    ProfScope ps(s, op == MPDX_RED_MIN ? "reduce_min" : (op == MPDX_RED_MAX ? "reduce_max" : "reduce_sum"));
    red_field<R>(s->stream, s->red_part, s->red_out, (const R *)arr, op, s->L);
  }
  return red_result(s, out);
}

// max Courant number of the advector (solver_3d.hpp:390-398)
template <class R>
int courant_impl(mpdx_solver *s, double *cfl)
{
  const R *g = (const R *)s->GC[0];
  red_launch<RK_CFL, R>(s->stream, s->red_part, s->red_out, g, (const R *)s->GC[1],
                        (const R *)s->GC[s->L.nd == 3 ? 2 : 0], s->L);
  return red_result(s, cfl);
}

// scale_gc (solver_3d.hpp:411-421, solver_2d.hpp:315-324): faces -1/2 ..
// n-3/2 along d times the interior across, then the advector halos
template <class R, class G>
int scale_gc(mpdx_solver *s, double cur_dt, double old_dt)
{
  const Layout &L = s->L;
  for (int d = 0; d < L.nd; ++d) {
    int lo[3] = {0, 0, 0}, hi[3] = {L.n[0] - 1, L.n[1] - 1, L.nd == 3 ? L.n[2] - 1 : 0};
    lo[d] = -1;
    hi[d] = L.n[d] - 2;
    Box bx = make_box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
    blaunch(s->stream, bx, k_scale<R>, (R *)s->GC[d], L, bx, R(cur_dt / old_dt));
  }
  PerPass<R, G> pp(s);
  R *Gb[3];
  R **gc = pp.av_of(s->GC, Gb);
  pp.xchng_vctr_alng(gc, 0);
  pp.xchng_vctr_nrml(gc, L.hw - 1);
  HIPCHK(hipGetLastError());
  return 0;
}

// hook_ante_loop chain (solver_common.hpp:191-211, solver_3d.hpp:375-388,
// mpdata_osc_3d.hpp:40-50)
template <class R, class G>
int ante_loop(mpdx_solver *s)
{
  PerPass<R, G> pp(s);
  R *Gb[3];
  R **gc = pp.av_of(s->GC, Gb);
  pp.xchng_vctr_alng(gc, 0);
  // var_dt: one rescaling for the initial advector (solver_common.hpp:199-210)
  if (s->p.var_dt) {
    double cfl = 0;
    if (int rc = courant_impl<R>(s, &cfl)) return rc;
    if (cfl > 0) {
      const double prev_dt = s->p.dt;
      s->p.dt *= s->p.max_courant / cfl;
      if (int rc = scale_gc<R, G>(s, s->p.dt, prev_dt)) return rc;
    }
  }
  red_launch<RK_DIV, R>(s->stream, s->red_part, s->red_out, (const R *)gc[0], (const R *)gc[1],
                        (const R *)gc[s->L.nd == 3 ? 2 : 0], s->L);
  double mx = 0;
  if (int rc = red_result(s, &mx)) return rc;
  const double eps = s->p.max_abs_div_eps < 0 ? (double)lim<R>::eps() : s->p.max_abs_div_eps;
  if ((R)mx > (R)eps) {
    s->err = "libmpdata++: initial advector field is divergent";
    return MPDX_EDIVERGENT;
  }
  pp.xchng_vctr_nrml(gc, s->L.hw - 1);
  return 0;
}

// ----------------------------------------------------------------- fused stepper

// rows next to an x or y edge whose z halo the boundary shell reads when
// z-boundary tiles compute the rest
constexpr int ZBAND = 5;

template <class R, class G>
struct Fused {
  mpdx_solver *s;
  PerPass<R, G> pp;
  const Layout &L;
  BcMap m;
  unsigned o;
  // With an interior kernel its tiles also compute the z-boundary cells,
  // and the boundary shell has no z slabs, for open bounds on both z sides
  // (zb) and for cyclic ones whose fields repeat plane z = 0 at z = n-1
  // (seam, decided for each advance call); with cyclic bounds everywhere
  // and fields repeating across every period (full) the tiles compute every
  // cell and there is no boundary shell at all
  bool zb, seam, full;

  explicit Fused(mpdx_solver *s_)
      : s(s_), pp(s_), L(s_->L), m(make_bcmap(s_)), o(s_->p.opts),
        zb(s_->use_zbnd && s_->L.nd == 3 && s_->p.bc[2][0] == MPDX_BC_OPEN && s_->p.bc[2][1] == MPDX_BC_OPEN),
        seam(s_->seam >= 1 && s_->L.nd == 3 && s_->p.bc[2][0] == MPDX_BC_CYCLIC && s_->p.bc[2][1] == MPDX_BC_CYCLIC),
        full((seam || s_->L.nd == 2) && s_->seam == 2 && s_->p.bc[0][0] == MPDX_BC_CYCLIC &&
             s_->p.bc[0][1] == MPDX_BC_CYCLIC && s_->p.bc[1][0] == MPDX_BC_CYCLIC &&
             s_->p.bc[1][1] == MPDX_BC_CYCLIC)
  {
  }

  bool ztiles() const { return zb || seam; }

  R *A(void *p) const { return static_cast<R *>(p); }

  dim3 grid_cells(int extra) const
  {
    const int inner = L.n[L.nd - 1] + extra;
    const unsigned gx = (inner + 255) / 256;
    if (L.nd == 3) return dim3(gx, L.n[1] + extra, L.n[0] + extra);
    return dim3(gx, L.n[0] + extra, 1);
  }

  // materialize the scalar halo of psi (the reference's xchng_sclr before
  // every pass, mpdata_osc_3d.hpp:73-202 via solver_common.hpp:127-136), all
  // slabs in one launch.  zbands: with z-boundary tiles the z halo is read
  // only by the boundary shell, i.e. in rows within ZBAND of an x or y edge.
  void psi_halo(R *a, bool zbands = false)
  {
    ProfScope ps(s, "fused_psi_halo");
    const int hw = L.hw;
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.nd == 3 ? L.n[2] : 1;
    const int zl = L.nd == 3 ? -hw : 0, zh = L.nd == 3 ? n2 - 1 + hw : 0;
    std::vector<std::pair<Box, int>> v;
    v.emplace_back(make_box(-hw, -1, -hw, n1 - 1 + hw, zl, zh), 0);
    v.emplace_back(make_box(n0, n0 - 1 + hw, -hw, n1 - 1 + hw, zl, zh), 0);
    v.emplace_back(make_box(0, n0 - 1, -hw, -1, zl, zh), 0);
    v.emplace_back(make_box(0, n0 - 1, n1, n1 - 1 + hw, zl, zh), 0);
    if (L.nd == 3) {
      const int zr[2][2] = {{-hw, -1}, {n2, n2 - 1 + hw}};
      for (const auto &r : zr) {
        if (!zbands) {
          v.emplace_back(make_box(0, n0 - 1, 0, n1 - 1, r[0], r[1]), 0);
        } else {
          for (const Box &b : zband_rows(r[0], r[1])) v.emplace_back(b, 0);
        }
      }
    }
    if (s->rec) {
      RecGroup grp(s);
      for (const auto &bt : v) {
        HaloTask<R> t{};
        t.a = a; t.L = L; t.m = m; t.bx = bt.first;
        record(s, t, box_grid(bt.first));
      }
      return;
    }
    boxlist_launch(v, [&](const BoxList &bl) {
      hipLaunchKernelGGL(k_fz_halo<R>, dim3(boxlist_blocks(bl)), dim3(256), 0, s->stream, a, L, m, bl);
    });
  }

  // rows within ZBAND of an x or y edge (x bands over every y, y bands
  // between them), over the z range [z0, z1]
  std::vector<Box> zband_rows(int z0, int z1) const
  {
    const int n0 = L.n[0], n1 = L.n[1];
    std::vector<Box> v;
    v.push_back(make_box(0, ZBAND - 1, 0, n1 - 1, z0, z1));
    v.push_back(make_box(n0 - ZBAND, n0 - 1, 0, n1 - 1, z0, z1));
    v.push_back(make_box(ZBAND, n0 - 1 - ZBAND, 0, ZBAND - 1, z0, z1));
    v.push_back(make_box(ZBAND, n0 - 1 - ZBAND, n1 - ZBAND, n1 - 1, z0, z1));
    return v;
  }

  // donor-cell update of the cells of the boxes, one launch per box (one
  // launch over several slabs measured slower next to the interior kernel)
  void donor_boxes(R *out, R *in, R **C, R in_scale, R out_scale, int flux_is_C, const std::vector<Box> &bxs)
  {
    R *c2 = L.nd == 3 ? C[2] : C[0];
    const int khn = (o & O_KHN) ? 1 : 0;
    RecGroup grp(s);
    for (const Box &bx : bxs) {
      if (bx.total <= 0) continue;
      const dim3 g = box_grid(bx);
      if (s->rec) {
        auto fill = [&](auto t) {
          t.out = out; t.in = in; t.C0 = C[0]; t.C1 = C[1]; t.C2 = c2; t.L = L;
          t.in_scale = in_scale; t.out_scale = out_scale; t.flux_is_C = flux_is_C; t.khn = khn; t.bx = bx;
          record(s, t, g);
        };
        if (o & O_NPA) fill(DonorTask<R, true>{});
        else fill(DonorTask<R, false>{});
        continue;
      }
      if (o & O_NPA)
        hipLaunchKernelGGL((k_fz_donor<R, true>), g, dim3(256), 0, s->stream, out, (const R *)in,
                           (const R *)C[0], (const R *)C[1], (const R *)c2, L, in_scale, out_scale, flux_is_C, khn,
                           bx);
      else
        hipLaunchKernelGGL((k_fz_donor<R, false>), g, dim3(256), 0, s->stream, out, (const R *)in,
                           (const R *)C[0], (const R *)C[1], (const R *)c2, L, in_scale, out_scale, flux_is_C, khn,
                           bx);
    }
  }

  void donor(R *out, R *in, R **C, R in_scale, R out_scale, int flux_is_C)
  {
    ProfScope ps(s, "fused_donor");
    donor_boxes(out, in, C, in_scale, out_scale, flux_is_C, {full_box(0)});
  }

  // the antidiffusive velocity for the owner cells of the boxes, one launch
  // per box
  template <bool IGA, bool ABS, bool PFC>
  void antidiff_t(R **V, R *psi, R **U, const std::vector<Box> &bxs)
  {
    RecGroup grp(s);
    for (const Box &bx : bxs) {
      if (bx.total <= 0) continue;
      const dim3 g = box_grid(bx);
      if (s->rec) {
        auto fill = [&](auto t, int nd) {
          t.V0 = V[0]; t.V1 = V[1]; t.V2 = nd == 3 ? V[2] : V[0]; t.psi = psi;
          t.U0 = U[0]; t.U1 = U[1]; t.U2 = nd == 3 ? U[2] : U[0]; t.L = L; t.bx = bx;
          record(s, t, g);
        };
        if (L.nd == 3) fill(AntidiffTask<R, G, 3, IGA, ABS, PFC>{}, 3);
        else fill(AntidiffTask<R, G, 2, IGA, ABS, PFC>{}, 2);
        continue;
      }
      if (L.nd == 3)
        hipLaunchKernelGGL((k_fz_antidiff<R, G, 3, IGA, ABS, PFC>), g, dim3(256), 0, s->stream,
                           V[0], V[1], V[2], (const R *)psi, (const R *)U[0], (const R *)U[1],
                           (const R *)U[2], L, bx);
      else
        hipLaunchKernelGGL((k_fz_antidiff<R, G, 2, IGA, ABS, PFC>), g, dim3(256), 0, s->stream,
                           V[0], V[1], V[0], (const R *)psi, (const R *)U[0], (const R *)U[1],
                           (const R *)U[0], L, bx);
    }
  }

  void antidiff_boxes(R **V, R *psi, R **U, const std::vector<Box> &bxs)
  {
    const bool iga = o & O_IGA, ab = o & O_ABS, pfc = o & O_PFC;
    if (iga) antidiff_t<true, false, false>(V, psi, U, bxs);
    else if (ab && pfc) antidiff_t<false, true, true>(V, psi, U, bxs);
    else if (ab) antidiff_t<false, true, false>(V, psi, U, bxs);
    else if (pfc) antidiff_t<false, false, true>(V, psi, U, bxs);
    else antidiff_t<false, false, false>(V, psi, U, bxs);
  }

  template <int ND, int SHAPE, bool NPA, bool ABS, bool IGA>
  void fct_s(R *out, R *psi, R *psin, R **V, R in_scale, R out_scale, const Box &bx)
  {
    const int khn = (o & O_KHN) ? 1 : 0;
    constexpr int TX = FzBrick<ND, SHAPE>::TX, TY = FzBrick<ND, SHAPE>::TY, TZ = FzBrick<ND, SHAPE>::TZ;
    dim3 g(ND == 3 ? (bx.ext[2] + TZ - 1) / TZ : 1, (bx.ext[1] + TY - 1) / TY, (bx.ext[0] + TX - 1) / TX);
    if (s->rec) {
      FctTask<R, G, ND, SHAPE, NPA, ABS, IGA> t{};
      t.out = out; t.psi = psi; t.psin = psin; t.V0 = V[0]; t.V1 = V[1]; t.V2 = V[ND == 3 ? 2 : 0];
      t.L = L; t.in_scale = in_scale; t.out_scale = out_scale; t.khn = khn; t.bx = bx;
      record(s, t, g);
      return;
    }
    hipLaunchKernelGGL((k_fz_fct<R, G, ND, SHAPE, NPA, ABS, IGA>), g, dim3(TX * TY * TZ), 0, s->stream, out,
                       (const R *)psi, (const R *)psin, (const R *)V[0], (const R *)V[1],
                       (const R *)V[ND == 3 ? 2 : 0], L, in_scale, out_scale, khn, bx);
  }

  // brick oriented along the box's thinnest dimension
  template <int ND, bool NPA, bool ABS, bool IGA>
  void fct_t(R *out, R *psi, R *psin, R **V, R in_scale, R out_scale, const Box &bx)
  {
    if (bx.total <= 0) return;
    int thin = 0;
    for (int d = 1; d < ND; ++d)
      if (bx.ext[d] < bx.ext[thin]) thin = d;
    if (thin == 0 || bx.ext[thin] > 8) fct_s<ND, 0, NPA, ABS, IGA>(out, psi, psin, V, in_scale, out_scale, bx);
    else if (thin == 1) fct_s<ND, 1, NPA, ABS, IGA>(out, psi, psin, V, in_scale, out_scale, bx);
    else fct_s<ND, 2, NPA, ABS, IGA>(out, psi, psin, V, in_scale, out_scale, bx);
  }

  template <int ND>
  void fct_nd(R *out, R *psi, R *psin, R **V, R in_scale, R out_scale, const Box &bx)
  {
    const bool npa = o & O_NPA, ab = o & O_ABS, iga = o & O_IGA;
    if (iga) { npa ? fct_t<ND, true, false, true>(out, psi, psin, V, in_scale, out_scale, bx)
                   : fct_t<ND, false, false, true>(out, psi, psin, V, in_scale, out_scale, bx); }
    else if (ab) { npa ? fct_t<ND, true, true, false>(out, psi, psin, V, in_scale, out_scale, bx)
                       : fct_t<ND, false, true, false>(out, psi, psin, V, in_scale, out_scale, bx); }
    else { npa ? fct_t<ND, true, false, false>(out, psi, psin, V, in_scale, out_scale, bx)
               : fct_t<ND, false, false, false>(out, psi, psin, V, in_scale, out_scale, bx); }
  }

  void fct_box(R *out, R *psi, R *psin, R **V, R in_scale, R out_scale, const Box &bx)
  {
    if (L.nd == 3) fct_nd<3>(out, psi, psin, V, in_scale, out_scale, bx);
    else fct_nd<2>(out, psi, psin, V, in_scale, out_scale, bx);
  }

  // the boxes of [f, l]^nd minus [f+wl, l-wh]^nd (faces x, then y, then z
  // slabs; the z slabs only without zb)
  std::vector<Box> shell(int f, int wl, int wh, bool zb_tiles = false) const
  {
    std::vector<Box> out;
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = d < L.nd ? f : 0;
      hi[d] = d < L.nd ? L.n[d] - 1 : 0;
    }
    int ilo[3], ihi[3];
    for (int d = 0; d < 3; ++d) {
      ilo[d] = d < L.nd ? lo[d] + wl : 0;
      ihi[d] = d < L.nd ? hi[d] - wh : 0;
    }
    bool empty = false;
    for (int d = 0; d < L.nd; ++d)
      if (ilo[d] > ihi[d]) empty = true;
    if (empty) {
      out.push_back(make_box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]));
      return out;
    }
    int cl[3] = {lo[0], lo[1], lo[2]}, ch[3] = {hi[0], hi[1], hi[2]};
    for (int d = 0; d < (zb_tiles && ztiles() ? 2 : L.nd); ++d) {
      int a[3] = {cl[0], cl[1], cl[2]}, b[3] = {ch[0], ch[1], ch[2]};
      b[d] = ilo[d] - 1;
      out.push_back(make_box(a[0], b[0], a[1], b[1], a[2], b[2]));
      a[d] = ihi[d] + 1;
      b[d] = ch[d];
      out.push_back(make_box(a[0], b[0], a[1], b[1], a[2], b[2]));
      cl[d] = ilo[d];
      ch[d] = ihi[d];
    }
    return out;
  }

  Box full_box(int f) const
  {
    return make_box(f, L.n[0] - 1, f, L.n[1] - 1, L.nd == 3 ? f : 0, L.nd == 3 ? L.n[2] - 1 : 0);
  }

  // the interior tile height: the solver's choice for the FP32 default-guard
  // kernels (16 or 32 rows for every pass; 0: 32 for the first corrective
  // pass, 24 for the later ones), the element type's default otherwise
  static constexpr bool kTileChoice = std::is_same<R, float>::value && std::is_same<G, GuardFast>::value;
  int tile_y() const
  {
    if (kTileChoice && (s->tile_y == 16 || (s->tile_y == 0 && s->rec && s->rec_tile16))) return 16;
    return itile_y_default<R>();
  }

  template <bool FCT, bool IGA, bool STOREV, bool FUSE0>
  void interior_t(R *out, R *psi, R *psin, R **U, R **V, R in_scale, R out_scale)
  {
    if constexpr (kTileChoice) {
      if (tile_y() == 16) return interior_ty<FCT, IGA, STOREV, FUSE0, 16>(out, psi, psin, U, V, in_scale, out_scale);
      // a later corrective pass (no in-tile donor cell) fits 80 VGPRs and
      // 64 KB of LDS at 24 rows: two blocks per CU, 6 waves per SIMD, whose
      // barrier phases are independent; 7 % faster than one 32-row block
      // (512^3 FCT, 1.23 vs 1.32 ms).  The first pass would spill there.
      // Steps recorded for batched launches (ensembles, batched equations)
      // keep 32 rows.
      if constexpr (!FUSE0 && !IGA)
        if (s->tile_y == 0 && !s->rec)
          return interior_ty<FCT, IGA, STOREV, FUSE0, 24>(out, psi, psin, U, V, in_scale, out_scale);
      return interior_ty<FCT, IGA, STOREV, FUSE0, 32>(out, psi, psin, U, V, in_scale, out_scale);
    } else {
      return interior_ty<FCT, IGA, STOREV, FUSE0, itile_y_default<R, FUSE0, STOREV>()>(out, psi, psin, U, V, in_scale,
                                                                                      out_scale);
    }
  }

  template <bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV>
  void interior_ty(R *out, R *psi, R *psin, R **U, R **V, R in_scale, R out_scale)
  {
    InteriorArgs<R> a;
    a.out = out; a.psi = psi; a.psin = psin;
    a.U0 = U[0]; a.U1 = U[1]; a.U2 = U[2];
    a.V0 = V[0]; a.V1 = V[1]; a.V2 = V[2];
    a.L = L;
    // velocity of cells 2..3 and n-4..n-3 comes from the boundary shell's
    // antidiff (which may run at the same time), so the interior kernel
    // stores only cells 4..n-5
    for (int d = 0; d < 3; ++d) {
      a.lo[d] = 2; a.hi[d] = L.n[d] - 3;
      a.vlo[d] = 4; a.vhi[d] = L.n[d] - 5;
    }
    a.in_scale = in_scale; a.out_scale = out_scale;
    constexpr int TY = ITile<TYV>::Y, TZ = ITile<TYV>::Z;
    // zb / seam: the tiles cover every z cell (and store the velocity of
    // every z face of the domain, open: face -1/2 included); the last tile
    // is shifted back to end on the last halo column
    a.zb = zb ? 1 : (seam ? 2 : 0);
    a.zs_max = std::max(L.n[2] - (TZ - 4), 0);
    if (ztiles()) {
      a.lo[2] = 0; a.hi[2] = L.n[2] - 1;
      a.vlo[2] = 0; a.vhi[2] = L.n[2] - 1;
    }
    // full: every x and y cell too (the last y tile shifted back like z's)
    a.ys_max = std::max(L.n[1] - (TY - 4), 0);
    if (full)
      for (int d = 0; d < 2; ++d) {
        a.lo[d] = 0; a.hi[d] = L.n[d] - 1;
        a.vlo[d] = 0; a.vhi[d] = L.n[d] - 1;
      }
    a.nb[0] = (a.hi[2] - a.lo[2] + 1 + TZ - 5) / (TZ - 4);
    a.nb[1] = (a.hi[1] - a.lo[1] + 1 + TY - 5) / (TY - 4);
    // x chunk: as long as possible (the 4-plane ramp of each chunk is
    // overhead; long blocks also hold CUs the concurrent boundary shell
    // waits for) while leaving at least two tiles per CU.  Every block
    // marches one chunk, one block per CU at a time, so the launch takes
    // ceil(blocks / CUs) rounds; when that leaves the last of two or more
    // rounds less than a third full, a nearby chunk count with the smallest
    // rounds x (planes + 4-step ramp + ~8 steps of block start-up) is taken
    // instead (256^3 FP64: 4 chunks of 63 planes took 3 rounds, the last
    // holding 8 of 520 blocks; 5 of 51 take 3 full ones, 9 % faster)
    const int nxc = a.hi[0] - a.lo[0] + 1;
    int chunk = s->ix_chunk;
    // blocks resident at a time: one per CU, two for the FP32 24-row tiles
    const long slots = (long)s->n_cu * (interior_waves_per_eu<R, TYV>() == 6 ? 2 : 1);
    if (chunk <= 0) {
      const long tiles = (long)a.nb[0] * a.nb[1];
      chunk = 256;
      while (chunk > 16 && tiles * ((nxc + chunk - 1) / chunk) < 2L * slots) chunk /= 2;
      const int c0 = (nxc + chunk - 1) / chunk;
      const long blocks = tiles * c0, rounds = (blocks + slots - 1) / slots;
      if (rounds >= 2 && 3 * (blocks - (rounds - 1) * slots) < slots) {
        long best = -1;
        for (int c = std::max(1, c0 - 2); c <= std::min(nxc, c0 + 3); ++c) {
          const int len = (nxc + c - 1) / c;
          const long cost = (tiles * c + slots - 1) / slots * (len + 12);
          if (best < 0 || cost < best) {
            best = cost;
            chunk = len;
          }
        }
      }
    }
    a.chunk = chunk;
    a.nb[2] = (nxc + a.chunk - 1) / a.chunk;
    const int ntile = a.nb[0] * a.nb[1] * a.nb[2];
    const dim3 g(8 * ((ntile + 7) / 8));   // tiles are spread over the 8 XCDs in the kernel
    if (full) launch_interior<FCT, IGA, STOREV, FUSE0, TYV, true>(a, g);
    else launch_interior<FCT, IGA, STOREV, FUSE0, TYV, false>(a, g);
  }

  template <bool FCT, bool IGA, bool STOREV, bool FUSE0, int TYV, bool WXY>
  void launch_interior(const InteriorArgs<R> &a, dim3 g)
  {
    if (s->rec) {
      // FP32 default guards: the mixed-variant task, which merges with
      // every other member's interior pass into one batched launch
      if constexpr (std::is_same<R, float>::value && std::is_same<G, GuardFast>::value && (TYV == 32 || TYV == 16)
                    && interior_mix_variant(FCT, IGA, STOREV, FUSE0)) {
        if (s->mix_tasks) {
          InteriorMixTask<R, G, TYV, WXY> t{};
          t.a = a;
          t.variant = (int)FCT | (int)IGA << 1 | (int)STOREV << 2 | (int)FUSE0 << 3;
          record(s, t, g);
          return;
        }
      }
      InteriorTask<R, G, FCT, IGA, STOREV, FUSE0, TYV, WXY> t{};
      t.a = a;
      record(s, t, g);
      return;
    }
    hipLaunchKernelGGL((k_fz_interior<R, G, FCT, IGA, STOREV, FUSE0, TYV, WXY>), g,
                       dim3(ITile<TYV>::Z, ITile<TYV>::Y), 0, s->stream, a);
  }

  // the 2D pass kernel over the interior cells 2..n-3 (velocity stored for
  // 4..n-5, like the 3D kernel), or with full over every cell
  void interior2d(R *out, R *psi, R *psin, R **U, R **V, bool storev, bool fuse0, R in_scale, R out_scale)
  {
    Pass2dArgs<R> a{};
    a.out = out; a.psi = psi; a.psin = psin;
    a.U0 = U[0]; a.U1 = U[1]; a.V0 = V[0]; a.V1 = V[1];
    a.L = L;
    for (int d = 0; d < 2; ++d) {
      a.lo[d] = full ? 0 : 2; a.hi[d] = full ? L.n[d] - 1 : L.n[d] - 3;
      a.vlo[d] = full ? 0 : 4; a.vhi[d] = full ? L.n[d] - 1 : L.n[d] - 5;
    }
    a.nb[0] = (a.hi[1] - a.lo[1] + kP2T) / kP2T;
    a.nb[1] = (a.hi[0] - a.lo[0] + kP2T) / kP2T;
    a.in_scale = in_scale; a.out_scale = out_scale;
    a.fuse0 = fuse0; a.storev = storev; a.wrap = full;
    const dim3 g((unsigned)(a.nb[0] * a.nb[1]));
    const bool fct = o & O_FCT, iga = o & O_IGA;
    if (fct && iga) launch_pass2d<true, true>(a, g);
    else if (fct) launch_pass2d<true, false>(a, g);
    else if (iga) launch_pass2d<false, true>(a, g);
    else launch_pass2d<false, false>(a, g);
  }

  template <bool FCT, bool IGA>
  void launch_pass2d(const Pass2dArgs<R> &a, dim3 g)
  {
    if (s->rec) {
      if (s->mix_tasks) {   // merges with every other 2D member's pass
        Pass2dMixTask<R, G> t{};
        t.a = a;
        t.variant = (int)FCT | (int)IGA << 1;
        record(s, t, g);
        return;
      }
      Pass2dTask<R, G, FCT, IGA> t{};
      t.a = a;
      record(s, t, g);
      return;
    }
    hipLaunchKernelGGL((k_fz_pass2d<R, G, FCT, IGA>), g, dim3(kP2Threads), 0, s->stream, a);
  }

  // psi: psi^k, or unused with fuse0 (psi^1 computed from psin and U = GC)
  void interior(R *out, R *psi, R *psin, R **U, R **V, bool storev, bool fuse0, R in_scale, R out_scale)
  {
    if (L.nd == 2) {
      ProfScope ps(s, fuse0 ? "fused_first_2d" : "fused_interior_2d");
      interior2d(out, psi, psin, U, V, storev, fuse0, in_scale, out_scale);
      return;
    }
    ProfScope ps(s, fuse0 ? ((o & O_FCT) ? (storev ? "fused_first_fct_v" : "fused_first_fct")
                                          : (storev ? "fused_first_v" : "fused_first"))
                          : ((o & O_FCT) ? (storev ? "fused_interior_fct_v" : "fused_interior_fct")
                                         : (storev ? "fused_interior_v" : "fused_interior")));
    const bool fct = o & O_FCT, iga = o & O_IGA;
#define MPDX_INT(F, I, S, Z) \
  if (fct == F && iga == I && storev == S && fuse0 == Z) interior_t<F, I, S, Z>(out, psi, psin, U, V, in_scale, out_scale)
#define MPDX_INT2(F, I, S) MPDX_INT(F, I, S, false); else MPDX_INT(F, I, S, true)
    MPDX_INT2(true, false, true); else MPDX_INT2(true, false, false); else MPDX_INT2(true, true, true);
    else MPDX_INT2(true, true, false); else MPDX_INT2(false, false, true); else MPDX_INT2(false, false, false);
    else MPDX_INT2(false, true, true); else MPDX_INT2(false, true, false);
#undef MPDX_INT2
#undef MPDX_INT
  }

  // The boundary shell of a pass (halo fills, shell kernels) and the
  // interior kernel touch disjoint output cells and read nothing the other
  // writes, so with an interior kernel the shell runs on the auxiliary
  // stream next to it (fork / join through events; a captured step keeps
  // the same dependencies).  Kernel timing runs everything on one stream.
  bool concurrent() const
  {
    return s->aux && s->ev_fork && s->ev_join && s->use_concurrency && !s->prof_on && !s->rec;
  }

  void fork()
  {
    (void)hipEventRecord(s->ev_fork, s->stream);
    (void)hipStreamWaitEvent(s->aux, s->ev_fork, 0);
  }
  void join(hipStream_t main)
  {
    (void)hipEventRecord(s->ev_join, s->aux);
    (void)hipStreamWaitEvent(main, s->ev_join, 0);
  }
  // launches inside the scope go to the auxiliary stream
  struct OnAux {
    mpdx_solver *s;
    hipStream_t main;
    bool on;
    OnAux(mpdx_solver *s_, bool on_) : s(s_), main(s_->stream), on(on_) { if (on) s->stream = s->aux; }
    ~OnAux() { s->stream = main; }
  };

  // boundary part of corrective pass it (or the whole pass without an
  // interior kernel): psi^k halo, antidiffusive velocity, its halos, FCT /
  // donor cell
  void corrective_shell(R *out, R *psik, R *psin, int it, int passes, R in_scale, R os, bool use_int)
  {
    R *Ub[3], *Vb[3];
    R **U = pp.unco(it, Ub);
    R **V = pp.corr(it, Vb);
    const bool fct = o & O_FCT;
    // next to z-boundary tiles the shell reads the z halos only near the x
    // and y edges: of psi within ZBAND rows; of the velocity in the rows
    // the shell computes (open: the tiles compute the z halo of the other
    // rows; cyclic: the tiles read none, and the tile rows the next pass's
    // shell reads get theirs from ring_zfill after the tiles are done)
    const bool zt = use_int && ztiles();
    pp.zskip = zt;
    // open z: the tangential components' z halo is zero (the normal fill)
    // and nothing but that fill writes it outside the edge bands (the tiles
    // store z = 0..n-1, the shell only its bands), so once an array has been
    // filled in full the later fills refill only the bands; a recorded step
    // (batched launches) always fills in full
    auto zeroed = [&](const R *a) {
      return std::find(s->zhalo_zeroed.begin(), s->zhalo_zeroed.end(), (const void *)a) != s->zhalo_zeroed.end();
    };
    const bool open_zt = zt && !seam && !s->rec;
    pp.zskip_nrml = zt && (seam || (open_zt && zeroed(V[0]) && zeroed(V[1])));
    auto mark_zeroed = [&]() {
      if (open_zt)
        for (int d = 0; d < 2; ++d)
          if (!zeroed(V[d])) s->zhalo_zeroed.push_back(V[d]);
    };
    const int zs = seam ? 4 : ZBAND;
    pp.zskip_x = R_(zs, L.n[0] - 1 - zs);
    pp.zskip_y = R_(zs, L.n[1] - 1 - zs);
    psi_halo(psik, zt);
    {
      ProfScope ps(s, "fused_antidiff");
      if (use_int)
        antidiff_boxes(V, psik, U, shell(-1, 5, 4, true));
      else
        antidiff_boxes(V, psik, U, {full_box(-1)});
    }
    if (fct) {
      {
        ProfScope ps(s, "fused_halo");
        pp.xchng_vctr_alng(V, 1);
        pp.xchng_vctr_nrml(V, 0);
      }
      mark_zeroed();
      ProfScope ps(s, "fused_fct");
      if (use_int) {
        RecGroup grp(s);
        for (const Box &b : shell(0, 2, 2, true)) fct_box(out, psik, psin, V, in_scale, os, b);
      } else
        fct_box(out, psik, psin, V, in_scale, os, full_box(0));
    } else {
      if (it != passes - 1) {
        ProfScope ps(s, "fused_halo");
        pp.xchng_vctr_nrml(V, 0);
        mark_zeroed();
      }
      if (use_int) {
        ProfScope ps(s, "fused_donor_shell");
        donor_boxes(out, psik, V, R(1), os, (o & O_IGA) ? 1 : 0, shell(0, 2, 2, true));
      } else {
        donor(out, psik, V, R(1), os, (o & O_IGA) ? 1 : 0);
      }
    }
  }

  // interior part of corrective pass it; with fuse0 the kernel recomputes
  // psi^1 (= psik, present only on the boundary shell) from psin and GC
  void corrective_interior(R *out, R *psik, R *psin, int it, int passes, R in_scale, R os, bool fuse0)
  {
    R *Ub[3], *Vb[3];
    R **U = pp.unco(it, Ub);
    R **V = pp.corr(it, Vb);
    const bool storev = it != passes - 1;   // the next pass reads this velocity as GC_unco
    interior(out, psik, psin, U, V, storev, fuse0, in_scale, os);
  }

  // seam mode, after a pass whose tiles stored the velocity: the z halo of
  // the tile rows next to the boundary shell, which the next pass's shell
  // reads (the tiles themselves read the z columns' periodic images, and
  // the shell's own rows got their z halo from the shell's fills)
  void ring_zfill(int it)
  {
    R *Vb[3];
    R **V = pp.corr(it, Vb);
    RingFillTask<R> t{};
    t.v0 = V[0]; t.v1 = V[1]; t.v2 = V[2]; t.L = L;
    t.x0 = 4; t.x1 = L.n[0] - 5; t.y0 = 4; t.y1 = L.n[1] - 5;
    const dim3 g((unsigned)((t.threads() + 255) / 256));
    if (s->rec) {
      record(s, t, g);
      return;
    }
    ProfScope ps(s, "fused_halo");
    hipLaunchKernelGGL(k_task<RingFillTask<R>>, g, dim3(256), 0, s->stream, t);
  }

  // one time step of equation e (solver_common.hpp:127-136 with advop inlined)
  void step_eqn(int e)
  {
    const int n_iters = s->p.n_iters;
    const int hs = s->p.hint_scale[e];
    const R in_scale = hs > 0 ? R(1) / R(1 << hs) : (hs < 0 ? R(1 << -hs) : R(1));
    const R out_scale = hs > 0 ? R(1 << hs) : (hs < 0 ? R(1) / R(1 << -hs) : R(1));
    const int lev = s->lev[e];
    R *psin = A(s->psi[e][lev]);
    R *fin = A(s->psi[e][lev ^ 1]);
    R *tmp = n_iters > 1 ? A(s->fz->psi_tmp[e]) : fin;
    // upwind_filter_freq (mpdata_osc_3d.hpp:195-198): only the donor-cell pass
    const bool upwind_only = s->p.upwind_filter_freq > 0 && s->timestep % s->p.upwind_filter_freq == 0;
    const int passes = upwind_only ? 1 : n_iters;
    auto out_of = [&](int it) { return ((passes - 1 - it) % 2 == 0) ? fin : tmp; };
    R *Gb[3];
    R **gc = pp.av_of(s->GC, Gb);
    const bool use_int = interior_supported(s->p, L) || (s->use_pass2d && pass2d_supported(s->p, L));
    if (full && use_int && passes > 1) {
      // every cell in the interior kernel, reading halos at their periodic
      // images: no halo fills and no boundary shell
      for (int it = 1; it < passes; ++it)
        corrective_interior(out_of(it), out_of(it - 1), psin, it, passes, in_scale,
                            it == passes - 1 ? out_scale : R(1), it == 1);
      s->lev[e] ^= 1;
      return;
    }
    psi_halo(psin, ztiles() && use_int && passes > 1);
    if (passes == 1 || !use_int) {
      donor(out_of(0), psin, gc, in_scale, passes == 1 ? out_scale : R(1), 0);
      for (int it = 1; it < passes; ++it)
        corrective_shell(out_of(it), out_of(it - 1), psin, it, passes, in_scale,
                         it == passes - 1 ? out_scale : R(1), false);
    } else {
      // pass 0 is needed only on the boundary shell (cells < 5 from a
      // boundary): the first corrective pass's interior kernel computes its
      // own psi^1
      const bool conc = concurrent();
      const hipStream_t main = s->stream;
      for (int it = 1; it < passes; ++it) {
        const R os = it == passes - 1 ? out_scale : R(1);
        if (conc) fork();
        {
          OnAux g(s, conc);
          if (it == 1) {
            ProfScope ps(s, "fused_donor_shell");
            donor_boxes(out_of(0), psin, gc, in_scale, R(1), 0, shell(0, 5, 5, true));
          }
          corrective_shell(out_of(it), out_of(it - 1), psin, it, passes, in_scale, os, true);
        }
        corrective_interior(out_of(it), out_of(it - 1), psin, it, passes, in_scale, os, it == 1);
        if (conc) join(main);
        if (seam && it != passes - 1) ring_zfill(it);
      }
    }
    s->lev[e] ^= 1;
  }
};

template <class R, class G>
int fused_step(mpdx_solver *s)
{
  Fused<R, G> fz(s);
  for (int e = 0; e < s->p.n_eqns; ++e) fz.step_eqn(e);
  return 0;
}

// time-varying advector for this step (solver_common.hpp:306-308), then the
// halo refill that the reference's calc_gc implementations perform; *changed
// says whether the advector changed (calc_gc's return value)
template <class R, class G>
int calc_gc_step(mpdx_solver *s, bool *changed_out = nullptr)
{
  if (changed_out) *changed_out = false;
  if (!s->calc_gc && s->gen_kind == MPDX_GC_NONE) return 0;
  const double t_half = s->time + 0.5 * s->p.dt;
  bool changed = false;
  if (s->gen_kind == MPDX_GC_DEFORM) {
    // amplitude from the table advance_impl uploaded for this call
    const Layout &L = s->L;
    Box bx = make_box(-1, L.n[0] - 1, -1, L.n[1] - 1, L.nd == 3 ? 0 : 0, L.nd == 3 ? L.n[2] - 1 : 0);
    ProfScope ps(s, "gen_deform");
    blaunch(s->stream, bx, k_gen_deform<R>, (R *)s->GC[0], (R *)s->GC[1],
           (R *)s->GC[L.nd == 3 ? 2 : 0], L, bx, (const double *)s->gen_amp, (const int *)s->gen_step,
           (const double *)s->gen_sx2, (const double *)s->gen_sy2);
    changed = true;
  }
  if (s->calc_gc) {
    const int r = s->calc_gc(s, t_half, s->calc_gc_user);
    if (r < 0) {
      if (s->err.empty()) s->err = "libmpdata++: calc_gc callback failed";
      return MPDX_EINVAL;
    }
    changed = changed || r != 0;
  }
  if (changed) {
    ProfScope ps(s, "gc_halo");
    PerPass<R, G> pp(s);
    R *Gb[3];
    R **gc = pp.av_of(s->GC, Gb);
    pp.xchng_vctr_alng(gc, 0);
    pp.xchng_vctr_nrml(gc, s->L.hw - 1);
  }
  HIPCHK(hipGetLastError());
  if (changed_out) *changed_out = changed;
  return 0;
}

// mpdata_rhs::update_rhs: zero the forcing over the interior, then the
// user's fn fills it (solvers/mpdata_rhs.hpp:52-70)
template <class R>
int update_rhs(mpdx_solver *s, double dt, int at)
{
  Box bi = interior_box(s->L);
  for (int e = 0; e < s->p.n_eqns; ++e)
    blaunch(s->stream, bi, k_set_box<R>, (R *)s->rhs[e], s->L, bi, R(0));
  HIPCHK(hipGetLastError());
  if (s->rhs_fn && s->rhs_fn(s, dt, at, s->rhs_user) < 0) {
    if (s->err.empty()) s->err = "libmpdata++: update_rhs callback failed";
    return MPDX_EINVAL;
  }
  return 0;
}

// mpdata_rhs::apply_rhs: state += dt_arg * R over the interior (:72-84)
template <class R>
void apply_rhs(mpdx_solver *s, double dt_arg)
{
  Box bi = interior_box(s->L);
  ProfScope ps(s, "rhs_apply");
  for (int e = 0; e < s->p.n_eqns; ++e)
    blaunch(s->stream, bi, k_axpy_box<R>, (R *)s->psi[e][s->lev[e]], (const R *)s->rhs[e], s->L, bi, R(dt_arg));
}

// ----------------------------------------------------------------- batched launches

// Recorded launches of several independent task sequences (the equations of
// one solver, or the members of an ensemble) merged into batched launches
// of k_task_list: the tasks of each batch and their block prefix table live
// in one device buffer.
void free_batch_plan(BatchPlan &p)
{
  if (p.dev) (void)hipFree(p.dev);
  p = BatchPlan();
}

// The merge: every sequence's recorded step (a member's, an equation's) is
// a sequence of groups of mutually independent launches.  Each round takes the kernel that heads
// the most members' current groups and launches all of its pending tasks in
// those groups as one batch; a member moves to its next group once its
// current one is done.  Each member's groups therefore run in order, and a
// batch holds tasks of different members (or independent tasks of one).
void schedule(const std::vector<TaskRecorder> &recs, std::vector<std::vector<const RecOp *>> &out,
              std::vector<std::vector<int>> &who)
{
  const size_t n = recs.size();
  std::vector<size_t> g(n, 0);
  std::vector<std::vector<char>> taken(n);
  auto settle = [&](size_t i) {
    while (g[i] < recs[i].groups.size()) {
      const auto &grp = recs[i].groups[g[i]];
      taken[i].resize(grp.size(), 0);
      bool done = true;
      for (size_t q = 0; q < grp.size(); ++q) done = done && taken[i][q];
      if (!done) return;
      ++g[i];
      taken[i].clear();
    }
  };
  for (size_t i = 0; i < n; ++i) settle(i);
  for (;;) {
    std::vector<std::pair<TaskLaunchFn, int>> count;   // kernel, members whose head holds it
    for (size_t i = 0; i < n; ++i) {
      if (g[i] >= recs[i].groups.size()) continue;
      const auto &grp = recs[i].groups[g[i]];
      std::vector<TaskLaunchFn> seen;
      for (size_t q = 0; q < grp.size(); ++q) {
        if (taken[i][q] || std::find(seen.begin(), seen.end(), grp[q].fn) != seen.end()) continue;
        seen.push_back(grp[q].fn);
        auto it = std::find_if(count.begin(), count.end(), [&](const auto &c) { return c.first == grp[q].fn; });
        if (it == count.end()) count.emplace_back(grp[q].fn, 1);
        else ++it->second;
      }
    }
    if (count.empty()) break;
    TaskLaunchFn best = count[0].first;
    int most = count[0].second;
    for (const auto &c : count)
      if (c.second > most) { best = c.first; most = c.second; }
    std::vector<const RecOp *> batch;
    std::vector<int> seqs;
    for (size_t i = 0; i < n; ++i) {
      if (g[i] >= recs[i].groups.size()) continue;
      const auto &grp = recs[i].groups[g[i]];
      for (size_t q = 0; q < grp.size(); ++q)
        if (!taken[i][q] && grp[q].fn == best) {
          batch.push_back(&grp[q]);
          taken[i][q] = 1;
          if (seqs.empty() || seqs.back() != (int)i) seqs.push_back((int)i);
        }
      settle(i);
    }
    out.push_back(std::move(batch));
    who.push_back(std::move(seqs));
  }
}

// merges the recorded sequences into batches and uploads their tasks and
// prefix tables (synchronously; not while a stream is being captured)
int pack_plan(const std::vector<TaskRecorder> &recs, BatchPlan &p, std::string &err)
{
  std::vector<std::vector<const RecOp *>> batches;
  std::vector<std::vector<int>> who;
  schedule(recs, batches, who);
  std::vector<int> last_of(recs.size(), -1);   // last batch holding tasks of each sequence
  // per batch its tasks, then its prefix table, each 256-byte aligned
  std::vector<unsigned char> blob;
  auto align = [&]() { blob.resize((blob.size() + 255) / 256 * 256); };
  for (const auto &b : batches) {
    EnsBatch eb;
    eb.fn = b[0]->fn;
    eb.n = (int)b.size();
    align();
    eb.toff = blob.size();
    for (const RecOp *op : b) blob.insert(blob.end(), op->bytes.begin(), op->bytes.end());
    align();
    eb.foff = blob.size();
    unsigned first = 0;
    for (const RecOp *op : b) {
      const unsigned char *f = reinterpret_cast<const unsigned char *>(&first);
      blob.insert(blob.end(), f, f + sizeof(unsigned));
      first += op->blocks;
    }
    eb.blocks = first;
    const int j = (int)p.batches.size();
    for (int i : who[j]) {
      if (last_of[i] >= 0 && std::find(eb.deps.begin(), eb.deps.end(), last_of[i]) == eb.deps.end())
        eb.deps.push_back(last_of[i]);
      last_of[i] = j;
    }
    p.batches.push_back(eb);
    p.tasks += b.size();
  }
  if (blob.empty()) blob.resize(256);
  hipError_t e = hipMalloc(&p.dev, blob.size());
  if (e == hipSuccess) e = hipMemcpy(p.dev, blob.data(), blob.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    err = std::string("libmpdata++: HIP error: ") + hipGetErrorString(e) + " (uploading batched tasks)";
    return MPDX_EHIP;
  }
  return 0;
}

void launch_plan(const BatchPlan &p, hipStream_t st)
{
  char *dev = static_cast<char *>(p.dev);
  for (const EnsBatch &b : p.batches)
    b.fn(st, dev + b.toff, reinterpret_cast<const unsigned *>(dev + b.foff), b.n, b.blocks);
}

// The same batches spread over st and the streams aux[0..naux-1], for
// capture into one graph: a batch follows only the batches it depends on
// (its sequences' earlier batches), so batches of unrelated sequences (2D
// and 3D members, members with other options) run side by side and fill
// each other's last block rounds.  A batch goes to the stream whose last
// batch it depends on, else to the stream idle longest; dependencies on
// other streams become event waits.  ev holds batches.size() + 1 events
// (one per batch and the fork).  Every auxiliary stream that gets a batch
// forks from st and joins it again at the end.
hipError_t launch_plan_streams(const BatchPlan &p, hipStream_t st, const hipStream_t *aux, int naux,
                               const hipEvent_t *ev)
{
  const int nb = (int)p.batches.size(), ns = naux + 1;
  auto S = [&](int k) { return k == 0 ? st : aux[k - 1]; };
  std::vector<int> where(nb, -1), last(ns, -1), used(ns, 0);
  for (int j = 0; j < nb; ++j) {
    int k = -1;
    for (int d : p.batches[j].deps)
      if (last[where[d]] == d) { k = where[d]; break; }
    if (k < 0) {
      k = 0;
      for (int q = 1; q < ns; ++q)
        if (last[q] < last[k]) k = q;
    }
    where[j] = k;
    last[k] = j;
    used[k] = 1;
  }
  char *dev = static_cast<char *>(p.dev);
  hipError_t e = hipEventRecord(ev[nb], st);
  for (int k = 1; k < ns && e == hipSuccess; ++k)
    if (used[k]) e = hipStreamWaitEvent(S(k), ev[nb], 0);
  std::fill(last.begin(), last.end(), -1);
  for (int j = 0; j < nb && e == hipSuccess; ++j) {
    const EnsBatch &b = p.batches[j];
    const int k = where[j];
    for (int d : b.deps)
      if (where[d] != k && e == hipSuccess) e = hipStreamWaitEvent(S(k), ev[d], 0);
    if (e != hipSuccess) break;
    b.fn(S(k), dev + b.toff, reinterpret_cast<const unsigned *>(dev + b.foff), b.n, b.blocks);
    e = hipEventRecord(ev[j], S(k));
    last[k] = j;
  }
  for (int k = 1; k < ns && e == hipSuccess; ++k)
    if (last[k] >= 0) e = hipStreamWaitEvent(st, ev[last[k]], 0);
  return e;
}

// the user's hooks of the mixed scheme (mpdata_rhs.hpp:112-124): they
// implement it with mpdx_update_rhs / mpdx_apply_rhs
int rhs_hook(mpdx_solver *s, int hook)
{
  if (!s->rhs_hook) {
    s->err = "libmpdata++: the mixed rhs scheme needs hooks (mpdx_set_rhs_hooks)";
    return MPDX_EINVAL;
  }
  if (s->rhs_hook(s, hook, s->rhs_hook_user) < 0) {
    if (s->err.empty()) s->err = "libmpdata++: rhs hook failed";
    return MPDX_EINVAL;
  }
  return 0;
}

// Seam mode for this advance call (1: z, 2: x, y and z): cyclic bounds on
// both sides of those dimensions, and every field the interior tiles read
// repeats plane 0 at plane n-1 along each of them (psi of every equation)
// and holds the periodic images in its halo (the advector, unless the
// on-device generator rewrites it each step: its corner tables repeat
// across the period by construction).  A step keeps this true: cells 0 and
// n-1 then see equal neighbourhoods and get equal updates.  A host calc_gc
// callback or forcing could break it, so those keep the slab path.
template <class R>
int seam_check_device(mpdx_solver *s, int *mode);

template <class R>
int seam_check(mpdx_solver *s, int *mode)
{
  if (s->seam_known && !s->seam_exposed) {
    *mode = s->seam;
    return 0;
  }
  if (int rc = seam_check_device<R>(s, mode)) return rc;
  s->seam_known = true;
  return 0;
}

template <class R>
int seam_check_device(mpdx_solver *s, int *mode)
{
  *mode = 0;
  const Layout &L = s->L;
  auto cyclic = [&](int d) { return s->p.bc[d][0] == MPDX_BC_CYCLIC && s->p.bc[d][1] == MPDX_BC_CYCLIC; };
  const int nd = L.nd;
  const bool xyz = cyclic(0) && cyclic(1);
  // 3D: z cyclic for mode 1, x and y too for mode 2; 2D: x and y for mode 2
  if (!s->use_seam || (nd == 3 ? !cyclic(2) : !xyz)) return 0;
  const bool kernel = nd == 3 ? interior_supported(s->p, L) : s->use_pass2d && pass2d_supported(s->p, L);
  if (s->p.n_iters < 2 || !kernel || s->calc_gc || s->rhs_scheme != MPDX_RHS_NONE) return 0;
  if (!s->seam_flag) {
    void *p = nullptr;
    if (int rc = dev_alloc(s, &p, 9 * sizeof(int))) return rc;
    s->seam_flag = (int *)p;
  }
  HIPCHK(hipMemsetAsync(s->seam_flag, 0, 9 * sizeof(int), s->stream));
  // tag: equation e (psi) or 64 + c (advector component c); flags 3 d .. 3 d + 2
  auto differ = [&](const void *a, int d, bool halo, const PlanePairs &pp, int tag) {
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    // rows: the interior (psi) or the rows the tiles read, one halo row on
    // each side (the advector); 2D rows run along the other dimension only
    const int i0 = halo ? -1 : 0, j0 = halo && nd == 3 ? -1 : 0;
    const int ni = L.n[d1] + (halo ? 2 : 0), nj = nd == 3 ? L.n[d2] + (halo ? 2 : 0) : 1;
    const long t = (long)ni * nj * pp.n;
    hipLaunchKernelGGL(k_pairs_differ<R>, dim3((unsigned)((t + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                       (const R *)a, L, d, i0, ni, j0, nj, pp, tag, s->seam_flag + 3 * d);
  };
  for (int d = xyz ? 0 : 2; d < nd; ++d) {
    const int n = L.n[d];
    const PlanePairs edge{1, {0}, {n - 1}};
    for (int e = 0; e < s->p.n_eqns; ++e) differ(s->psi[e][s->lev[e]], d, false, edge, e);
    if (s->gen_kind != MPDX_GC_NONE) continue;
    // components tangential to d at cells (halo cells -m and n-1+m), the
    // normal one at faces + 1/2 (halo faces -m and n-2+m), m = 1..hw
    PlanePairs tang{1, {0}, {n - 1}}, nrml{1, {n - 1}, {0}};
    for (int m = 1; m <= L.hw; ++m) {
      tang.pa[tang.n] = -m; tang.pb[tang.n++] = n - 1 - m;
      tang.pa[tang.n] = n - 1 + m; tang.pb[tang.n++] = m;
      nrml.pa[nrml.n] = -m; nrml.pb[nrml.n++] = n - 1 - m;
      if (m >= 2) { nrml.pa[nrml.n] = n - 2 + m; nrml.pb[nrml.n++] = m - 1; }
    }
    for (int c = 0; c < nd; ++c) differ(s->GC[c], d, true, c == d ? nrml : tang, 64 + c);
  }
  HIPCHK(hipGetLastError());
  int h[9] = {};
  HIPCHK(hipMemcpyAsync(h, s->seam_flag, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const bool z_ok = nd == 2 || h[6] == 0, xy_ok = xyz && h[0] == 0 && h[3] == 0;
  *mode = !z_ok ? 0 : (xy_ok ? 2 : (nd == 3 ? 1 : 0));
  // MPDX_SEAM_DEBUG=1: say which field keeps which dimension off the seam path
  if (std::getenv("MPDX_SEAM_DEBUG"))
    for (int d = 0; d < 3; ++d)
      if (h[3 * d])
        std::fprintf(stderr, "mpdx seam check: dimension %d: %s %d, plane pair %d, row (%d, %d) differ\n", d,
                     h[3 * d] / 16 >= 64 ? "advector component" : "equation", h[3 * d] / 16 % 64,
                     h[3 * d] % 16 - 1, h[3 * d + 1], h[3 * d + 2]);
  return 0;
}

// ---------------------------------------------------------------- var_dt output

int out_call(mpdx_solver *s, int kind, double t)
{
  if (s->out_fn(s, kind, t, s->out_user) != 0) {
    if (s->err.empty()) s->err = "libmpdata++: output callback failed";
    return MPDX_EINVAL;
  }
  return 0;
}

// the current state of the output equations into interpolation slot k
template <class R>
int out_save(mpdx_solver *s, int k)
{
  for (int v = 0; v < s->out_n; ++v) {
    const int e = s->out_var[v];
    HIPCHK(hipMemcpyAsync(s->out_buf[k][v], s->psi[e][s->lev[e]], sizeof(R) * (size_t)s->L.size,
                          hipMemcpyDeviceToDevice, s->stream));
  }
  return 0;
}

// hook_ante_loop (output_common.hpp:47-62): the buffers take the state and
// the caller writes const.h5 and the first record
template <class R>
int out_start(mpdx_solver *s, double t_end)
{
  if (int rc = out_save<R>(s, 0)) return rc;
  s->out_cnt = 0;
  s->out_rec = s->time;
  return out_call(s, MPDX_OUT_START, t_end);
}

// hook_ante_step (output_common.hpp:72-101), once the step's dt is set: a
// step that reaches the next multiple of outfreq starts a record (its time
// k * outfreq, out_ord states to collect), and while one is open the state
// before the step is saved
template <class R>
int out_ante(mpdx_solver *s)
{
  const double next_time = s->time + s->p.dt;
  const long next_idx = (long)std::floor(next_time / s->out_freq);
  const long curr_idx = (long)std::floor(s->time / s->out_freq);
  s->out_cnt = s->out_cnt > 0 ? s->out_cnt - 1 : 0;
  if (next_idx > curr_idx || s->out_cnt > 0) {
    if (s->out_cnt == 0) {
      s->out_cnt = s->out_ord;
      s->out_rec = (double)next_idx * s->out_freq;
    }
    const int k = s->out_ord - s->out_cnt;
    s->out_t[k] = s->time;
    if (int rc = out_save<R>(s, k)) return rc;
  }
  return 0;
}

// hook_post_step (output_common.hpp:104-156): the step that completes a
// record interpolates it into slot 0 and hands it to the caller
template <class R>
int out_post(mpdx_solver *s)
{
  if (s->out_cnt != 1) return 0;
  const double t = s->out_rec, t0 = s->out_t[0];
  for (int v = 0; v < s->out_n; ++v) {
    const int e = s->out_var[v];
    R *y0 = (R *)s->out_buf[0][v];
    const R *psi = (const R *)s->psi[e][s->lev[e]];
    if (s->out_ord == 1) {
      const double t1 = s->time;
      launch(s->stream, s->L.size, k_out_lin<R>, y0, psi, s->L.size, R((t1 - t) / (t1 - t0)),
             R((t - t0) / (t1 - t0)));
    } else {
      const double t1 = s->out_t[1], t2 = s->time;
      launch(s->stream, s->L.size, k_out_quad<R>, y0, (const R *)s->out_buf[1][v], psi, s->L.size,
             R(t1 - t0), R(t2 - t0), R(t2 - t1), R(t - t0), R(t - t1));
    }
  }
  HIPCHK(hipGetLastError());
  return out_call(s, MPDX_OUT_RECORD, t);
}

// nt steps, or with by_time the reference's advance(real_t) of a var_dt
// solver (solver_common.hpp:292-355): steps while time < t_end or
// additional_steps > 0, which starts at out_intrp_ord (1, the default) and
// drops on each step ending at or past t_end, so at least one step is taken
// Batched equations without the switch: every cell in the interior tiles
// (all-cyclic seam mode, no boundary shell to overlap), and the velocity
// scratch of equations 1.. (n_iters - 1 sets of nd arrays each) within an
// eighth of the device's memory
template <class R>
bool auto_batch(mpdx_solver *s)
{
  if (!s->auto_batch_eqns || s->seam != 2 || s->L.nd != 3 || s->active_path != MPDX_PATH_FUSED) return false;
  if (s->auto_batch_fits < 0) {   // the memory test, once per solver
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      total_b = 0;
    }
    const double arr = (double)s->L.size * sizeof(R);
    const double extra = arr * s->L.nd * std::min(s->p.n_iters - 1, 2) * (s->p.n_eqns - 1);
    s->auto_batch_fits = total_b > 0 && extra <= (double)total_b / 8;
  }
  return s->auto_batch_fits == 1;
}

template <class R, class G>
int advance_impl(mpdx_solver *s, int64_t nt, double t_end = 0, bool by_time = false)
{
  const bool var_dt = s->p.var_dt;
  nt += s->timestep;
  if (s->timestep == 0) {
    if (int rc = ante_loop<R, G>(s)) return rc;
    // mpdata_rhs::hook_ante_loop (solvers/mpdata_rhs.hpp:126-145)
    if (s->rhs_scheme == MPDX_RHS_TRAPEZ)
      if (int rc = update_rhs<R>(s, s->p.dt / 2, 0)) return rc;
    if (s->rhs_scheme == MPDX_RHS_MIXED)
      if (int rc = rhs_hook(s, MPDX_HOOK_ANTE_LOOP)) return rc;
    if (s->out_fn)
      if (int rc = out_start<R>(s, by_time ? t_end : s->time + (double)(nt - s->timestep) * s->p.dt)) return rc;
  }
  const int rs = s->rhs_scheme;
  const bool fused = s->active_path == MPDX_PATH_FUSED;
  if (!fused)
    if (int rc = ensure_perpass(s)) return rc;
  int seam = 0;
  if (fused)
    if (int rc = seam_check<R>(s, &seam)) return rc;
  s->seam = seam;
  PerPass<R, G> pp(s);
  // on-device advector generator: the amplitude of every step of this call,
  // evaluated on the host like the oracle's, and a device step index
  std::vector<double> amps;
  const bool gen = s->gen_kind == MPDX_GC_DEFORM;
  if (gen) {
    const long n = nt - s->timestep;
    if (n > s->gen_amp_cap) {
      // a captured step reads the table by address: drop the graphs before
      // the table moves, and free the old one once nothing uses it
      HIPCHK(hipStreamSynchronize(s->stream));
      drop_step_graphs(s);
      const long cap = std::max(n, std::max(2 * s->gen_amp_cap, 64L));
      void *pa = nullptr;
      if (int rc = dev_alloc(s, &pa, sizeof(double) * (size_t)cap)) return rc;
      if (s->gen_amp) {
        for (auto &p : s->allocs)
          if (p == s->gen_amp) p = nullptr;
        (void)hipFree(s->gen_amp);
      }
      s->gen_amp = (double *)pa;
      s->gen_amp_cap = cap;
    }
    if (!s->gen_step) {
      void *pc = nullptr;
      if (int rc = dev_alloc(s, &pc, sizeof(int))) return rc;
      s->gen_step = (int *)pc;
    }
    amps.resize(n > 0 ? n : 1);
    for (long q = 0; q < n; ++q) {
      const double t_half = (s->timestep + q) * s->p.dt + 0.5 * s->p.dt;
      amps[q] = s->gen_scale * std::cos(3.141592653589793 * t_half / s->gen_T);
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(s->gen_amp, amps.data(), sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, s->stream, s->gen_step, 0);
  }
  // the fused step is replayed from a captured hipGraph unless something in
  // it changes from step to step in a way a graph cannot follow (host
  // calc_gc callback, upwind filter steps) or kernels are being timed
  // individually (the legacy null stream cannot be captured)
  const bool graphs = fused && s->use_graphs && !s->prof_on && !s->calc_gc && s->stream != nullptr &&
                      s->p.upwind_filter_freq <= 0 && rs == MPDX_RHS_NONE;
  // bounded queue: before step q the host waits for step q - kInflight, so
  // the panic flag is looked at while at most kInflight steps are queued
  if (!s->inflight[0])
    for (int q = 0; q < mpdx_solver::kInflight; ++q)
      if (hipEventCreateWithFlags(&s->inflight[q], hipEventDisableTiming) != hipSuccess) {
        s->inflight[q] = nullptr;
        (void)hipGetLastError();
      }
  int64_t queued = 0;
  bool stepped = false;
  while (by_time ? (s->time < t_end || !stepped) : s->timestep < nt) {
    stepped = true;
    if (queued >= mpdx_solver::kInflight) {
      hipEvent_t ev = s->inflight[queued % mpdx_solver::kInflight];
      if (ev) HIPCHK(hipEventSynchronize(ev));
    }
    if (panicked(s)) break;
    int adv = 1;   // steps this iteration advances
    if (graphs) {
      const bool batched = s->p.n_eqns > 1 && (s->batch_eqns || auto_batch<R>(s));
      if (s->graph_steps == 0) {
        long cells = (long)s->p.n_eqns;
        for (int d = 0; d < s->L.nd; ++d) cells *= s->L.n[d];
        s->graph_steps = cells <= (1L << 21) ? 8 : 1;
      }
      const int gs = !var_dt && !batched && nt - s->timestep >= s->graph_steps ? s->graph_steps : 1;
      const int key = (gs > 1 ? 6 : 0) + ((s->lev[0] & 1) | (s->seam << 1));
      if (!s->graph[key]) {
        int lev_save[MPDX_MAX_EQNS];
        for (int e = 0; e < s->p.n_eqns; ++e) lev_save[e] = s->lev[e];
        // several equations: the launches of each equation's step are
        // recorded and merged into batched launches (one per kernel and
        // round, covering every equation), uploaded before the capture
        if (batched) {
          // equations 1.. get velocity scratch of their own
          for (int e = 1; e < s->p.n_eqns; ++e)
            for (int d = 0; d < s->L.nd; ++d) {
              if (s->p.n_iters > 1)
                if (int rc = alloc_missing(s, &s->tmp_eq[e][0][d])) return rc;
              if (s->p.n_iters > 2)
                if (int rc = alloc_missing(s, &s->tmp_eq[e][1][d])) return rc;
            }
          std::vector<TaskRecorder> recs(s->p.n_eqns);
          // 16-row tiles unless the solver chose a height: the batched
          // passes hold n_eqns times the blocks (8 tracers, 512 x 512 x 128:
          // 7.8 ms per step against 8.3 at 32 rows)
          s->rec_tile16 = true;
          for (int e = 0; e < s->p.n_eqns; ++e) {
            s->rec = &recs[e];
            s->tmp_sel = e;
            Fused<R, G> fz(s);
            fz.step_eqn(e);
            s->rec = nullptr;
            s->tmp_sel = -1;
          }
          s->rec_tile16 = false;
          for (int e = 0; e < s->p.n_eqns; ++e) s->lev[e] = lev_save[e];
          if (s->eqplan[key].dev) (void)hipFree(s->eqplan[key].dev);
          s->eqplan[key] = BatchPlan();
          if (int rc = pack_plan(recs, s->eqplan[key], s->err)) return rc;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        for (int q = 0; q < gs && !rc; ++q) {
          rc = calc_gc_step<R, G>(s);
          if (!rc) {
            if (batched) launch_plan(s->eqplan[key], s->stream);
            else rc = fused_step<R, G>(s);
          }
          if (gen) hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, s->stream, s->gen_step, -1);
        }
        hipError_t ec = hipStreamEndCapture(s->stream, &g);
        for (int e = 0; e < s->p.n_eqns; ++e) s->lev[e] = lev_save[e];
        if (rc) return rc;
        HIPCHK(ec);
        hipError_t ei = hipGraphInstantiate(&s->graph[key], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ei);
      }
      if (s->out_fn)
        if (int rc = out_ante<R>(s)) return rc;
      HIPCHK(hipGraphLaunch(s->graph[key], s->stream));
      for (int e = 0; e < s->p.n_eqns; ++e) s->lev[e] ^= gs & 1;
      adv = gs;
    } else {
      bool var_gc = false;
      if (int rc = calc_gc_step<R, G>(s, &var_gc)) return rc;
      // var_dt with a time-varying advector: shrink / grow dt until the
      // Courant number is at most max_courant, only in steps whose calc_gc
      // reported a changed advector (solver_common.hpp:306-324); the
      // comparisons are in real_t like the reference's
      if (var_dt && var_gc) {
        double cfl = 0;
        if (int rc = courant_impl<R>(s, &cfl)) return rc;
        int guard = 0;
        while ((R)cfl > R(0)) {
          s->p.dt *= s->p.max_courant / cfl;
          if (int rc = calc_gc_step<R, G>(s)) return rc;
          if (int rc = courant_impl<R>(s, &cfl)) return rc;
          if (!((R)cfl > (R)s->p.max_courant)) break;
          if (++guard == 64) {
            s->err = "libmpdata++: var_dt did not bring the Courant number below max_courant";
            return MPDX_EINVAL;
          }
        }
      }
      // mpdata_rhs::hook_ante_step (solvers/mpdata_rhs.hpp:135-160)
      if (rs == MPDX_RHS_EULER_A || rs == MPDX_RHS_EULER_B) {
        if (int rc = update_rhs<R>(s, s->p.dt, 0)) return rc;
        if (rs == MPDX_RHS_EULER_B) apply_rhs<R>(s, s->p.dt);
      } else if (rs == MPDX_RHS_TRAPEZ) {
        apply_rhs<R>(s, s->p.dt / 2);
      } else if (rs == MPDX_RHS_MIXED) {
        if (int rc = rhs_hook(s, MPDX_HOOK_ANTE_STEP)) return rc;
      }
      // output_common::hook_ante_step runs after its parents' (output<rhs<...>>)
      if (s->out_fn)
        if (int rc = out_ante<R>(s)) return rc;
      if (fused) {
        if (int rc = fused_step<R, G>(s)) return rc;
      } else {
        pp.step();
      }
      if (gen) hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, s->stream, s->gen_step, -1);
    }
    HIPCHK(hipGetLastError());
    if (hipEvent_t ev = s->inflight[queued % mpdx_solver::kInflight]) HIPCHK(hipEventRecord(ev, s->stream));
    ++queued;
    s->timestep += adv;
    s->time = var_dt ? s->time + s->p.dt : s->timestep * s->p.dt;
    // mpdata_rhs::hook_post_step (solvers/mpdata_rhs.hpp:162-183)
    if (rs == MPDX_RHS_EULER_A) {
      apply_rhs<R>(s, s->p.dt);
    } else if (rs == MPDX_RHS_TRAPEZ) {
      if (int rc = update_rhs<R>(s, s->p.dt / 2, 1)) return rc;
      apply_rhs<R>(s, s->p.dt / 2);
    } else if (rs == MPDX_RHS_MIXED) {
      if (int rc = rhs_hook(s, MPDX_HOOK_POST_STEP)) return rc;
    }
    if (s->out_fn)
      if (int rc = out_post<R>(s)) return rc;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

// element type / guard dispatch of the per-member work of an ensemble
int ante_dispatch(mpdx_solver *s)
{
  if (s->esz == 8) return ante_loop<double, double>(s);
  if (s->p.guard == MPDX_GUARD_NATIVE) return ante_loop<float, float>(s);
  if (s->p.guard == MPDX_GUARD_F64_EXACT) return ante_loop<float, double>(s);
  return ante_loop<float, GuardFast>(s);
}

int seam_dispatch(mpdx_solver *s)
{
  int mode = 0;
  int rc = 0;
  if (s->esz == 8) rc = seam_check<double>(s, &mode);
  else rc = seam_check<float>(s, &mode);
  s->seam = mode;
  return rc;
}

int step_dispatch(mpdx_solver *s)
{
  if (s->esz == 8) return fused_step<double, double>(s);
  if (s->p.guard == MPDX_GUARD_NATIVE) return fused_step<float, float>(s);
  if (s->p.guard == MPDX_GUARD_F64_EXACT) return fused_step<float, double>(s);
  return fused_step<float, GuardFast>(s);
}

}  // namespace

// ----------------------------------------------------------------- ensembles

// Independent solvers stepped together.  One time step of every member is
// recorded as tasks (the launches its own step would make, with its own
// arrays, boxes and scales), the tasks are merged across members into
// batched launches of k_task_list (one launch per kernel and round of the
// schedule below), and the batched launches are captured on the ensemble's
// stream as one linear hipGraph, replayed once per step.  A plan is keyed by
// the members' time levels; two alternate.
namespace {

struct EnsPlan {
  std::vector<int> key;              // members' time levels the plan was recorded at
  BatchPlan bp;
  hipGraphExec_t graph = nullptr;
  std::vector<hipEvent_t> ev;        // the capture's fork and per-batch events
};

}  // namespace

struct mpdx_ensemble {
  std::vector<mpdx_solver *> m;
  hipStream_t stream = nullptr;
  // streams that batches of unrelated members are spread over while a
  // step is captured (MPDX_ENS_STREAMS=n: streams in all; unset: 1, one
  // linear sequence)
  static constexpr int kAux = 7;
  hipStream_t aux[kAux] = {};
  int n_aux = 0;
  EnsPlan plan[2];
  int next_slot = 0;
  static constexpr int kInflight = 6;
  hipEvent_t inflight[kInflight] = {};
  std::string err;
};

namespace {

#define EHIPCHK(expr)                                                       \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      e->err = std::string("libmpdata++: HIP error: ") + hipGetErrorString(e_) \
               + " (" #expr ")";                                            \
      return MPDX_EHIP;                                                     \
    }                                                                       \
  } while (0)

// members' time levels and seam modes
std::vector<int> level_key(const mpdx_ensemble *e)
{
  std::vector<int> k;
  for (const mpdx_solver *s : e->m) {
    for (int q = 0; q < s->p.n_eqns; ++q) k.push_back(s->lev[q]);
    k.push_back(s->seam);
  }
  return k;
}

void free_plan(EnsPlan &p)
{
  if (p.graph) (void)hipGraphExecDestroy(p.graph);
  free_batch_plan(p.bp);
  for (hipEvent_t ev : p.ev)
    if (ev) (void)hipEventDestroy(ev);
  p = EnsPlan();
}

void drop_plans(mpdx_ensemble *e)
{
  for (EnsPlan &p : e->plan) free_plan(p);
  e->next_slot = 0;
}

int member_error(mpdx_ensemble *e, size_t i, const mpdx_solver *s, int rc)
{
  e->err = "libmpdata++: ensemble member " + std::to_string(i) + ": " + s->err;
  return rc;
}

// records one step of every member, merges it into batches, uploads the
// tasks and captures the batched launches as one graph
int build_plan(mpdx_ensemble *e, EnsPlan &p)
{
  const std::vector<int> saved = level_key(e);
  std::vector<TaskRecorder> recs(e->m.size());
  int rc = 0;
  size_t fail_at = 0;
  for (size_t i = 0; i < e->m.size() && !rc; ++i) {
    mpdx_solver *s = e->m[i];
    s->rec = &recs[i];
    rc = step_dispatch(s);
    s->rec = nullptr;
    fail_at = i;
  }
  size_t q = 0;
  for (mpdx_solver *s : e->m) {
    for (int r = 0; r < s->p.n_eqns; ++r) s->lev[r] = saved[q++];
    ++q;   // (the seam mode)
  }
  if (rc) return member_error(e, fail_at, e->m[fail_at], rc);

  if (int rc2 = pack_plan(recs, p.bp, e->err)) return rc2;
  if (e->n_aux > 0) {
    p.ev.assign(p.bp.batches.size() + 1, nullptr);
    for (hipEvent_t &ev : p.ev) EHIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  hipGraph_t graph = nullptr;
  EHIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  hipError_t el = hipSuccess;
  if (e->n_aux > 0)
    el = launch_plan_streams(p.bp, e->stream, e->aux, e->n_aux, p.ev.data());
  else
    launch_plan(p.bp, e->stream);
  if (el == hipSuccess) el = hipGetLastError();
  const hipError_t ec = hipStreamEndCapture(e->stream, &graph);
  if (el != hipSuccess && graph) {
    (void)hipGraphDestroy(graph);
    graph = nullptr;
  }
  EHIPCHK(el);
  EHIPCHK(ec);
  const hipError_t ei = hipGraphInstantiate(&p.graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  EHIPCHK(ei);
  p.key = saved;
  return 0;
}

// first-advance work of members that have not stepped yet (hook_ante_loop),
// each member's seam mode for this call, and every member's own pending
// work finished before the ensemble's stream touches its arrays
int ensemble_ante(mpdx_ensemble *e)
{
  for (size_t i = 0; i < e->m.size(); ++i) {
    mpdx_solver *s = e->m[i];
    if (s->timestep == 0)
      if (int rc = ante_dispatch(s)) return member_error(e, i, s, rc);
    if (int rc = seam_dispatch(s)) return member_error(e, i, s, rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) {
      e->err = "libmpdata++: ensemble member " + std::to_string(i) + ": HIP error";
      return MPDX_EHIP;
    }
  }
  return 0;
}

// launches one ensemble step (building the plan for the current levels when
// needed) and advances every member's clock
int ensemble_step(mpdx_ensemble *e)
{
  const std::vector<int> k = level_key(e);
  int slot = -1;
  for (int q = 0; q < 2; ++q)
    if (e->plan[q].graph && e->plan[q].key == k) slot = q;
  if (slot < 0) {
    slot = e->next_slot;
    e->next_slot ^= 1;
    EHIPCHK(hipStreamSynchronize(e->stream));   // the slot's old graph may still run
    free_plan(e->plan[slot]);
    if (int rc = build_plan(e, e->plan[slot])) {
      free_plan(e->plan[slot]);
      return rc;
    }
  }
  EHIPCHK(hipGraphLaunch(e->plan[slot].graph, e->stream));
  for (mpdx_solver *s : e->m) {
    for (int r = 0; r < s->p.n_eqns; ++r) s->lev[r] ^= 1;
    s->timestep++;
    s->time = s->timestep * s->p.dt;
  }
  return 0;
}

bool ensemble_panic(const mpdx_ensemble *e)
{
  for (const mpdx_solver *s : e->m)
    if (panicked(s)) return true;
  return false;
}

// nt steps with at most kInflight of them queued ahead of the host (panic
// flags are polled between steps)
int ensemble_run(mpdx_ensemble *e, int64_t nt)
{
  if (!e->inflight[0])
    for (hipEvent_t &ev : e->inflight)
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
        (void)hipGetLastError();
      }
  for (int64_t t = 0; t < nt; ++t) {
    if (t >= mpdx_ensemble::kInflight)
      if (hipEvent_t ev = e->inflight[t % mpdx_ensemble::kInflight]) EHIPCHK(hipEventSynchronize(ev));
    if (ensemble_panic(e)) break;
    if (int rc = ensemble_step(e)) return rc;
    if (hipEvent_t ev = e->inflight[t % mpdx_ensemble::kInflight]) EHIPCHK(hipEventRecord(ev, e->stream));
  }
  return 0;
}

}  // namespace

// ----------------------------------------------------------------- C ABI

// A null handle or output pointer is an error (MPDX_EINVAL) rather than a
// crash; the message is then available from mpdx_last_error(NULL) /
// mpdx_ensemble_last_error(NULL).
#define MPDX_NEED(cond, what)                                     \
  do {                                                            \
    if (!(cond)) {                                                \
      g_create_error = "libmpdata++: null " what;                 \
      return MPDX_EINVAL;                                         \
    }                                                             \
  } while (0)

extern "C" {

void mpdx_params_default(mpdx_params *p)
{
  std::memset(p, 0, sizeof(*p));
  p->n_dims = 3;
  p->real_bytes = MPDX_REAL_F64;
  p->opts = MPDX_OPT_IGA | MPDX_OPT_FCT;   // ct_params_default_t (opts.hpp:83)
  p->n_iters = 2;                           // mpdata_common.hpp:76
  p->n_eqns = 1;
  p->max_abs_div_eps = -1;
  p->guard = MPDX_GUARD_F64;
  p->path = MPDX_PATH_AUTO;
  p->max_courant = 0.5;                     // solver_common.hpp:218
}

int mpdx_create(const mpdx_params *p, mpdx_solver **out)
{
  MPDX_NEED(out, "output pointer");
  *out = nullptr;
  const uint32_t unsupported = MPDX_OPT_TOT | MPDX_OPT_NUG | MPDX_OPT_DFL | ~0x1FFu;
  auto fail = [&](int code, const char *msg) {
    g_create_error = msg;
    return code;
  };
  if (!p) return fail(MPDX_EINVAL, "libmpdata++: null parameters");
  if (p->n_dims != 2 && p->n_dims != 3) return fail(MPDX_EINVAL, "libmpdata++: n_dims must be 2 or 3");
  for (int d = 0; d < p->n_dims; ++d)
    if (p->grid[d] < 1) return fail(MPDX_EINVAL, "libmpdata++: bogus grid size");
  if (p->real_bytes != 4 && p->real_bytes != 8) return fail(MPDX_EINVAL, "libmpdata++: real_bytes must be 4 or 8");
  if (p->opts & unsupported) return fail(MPDX_EUNSUPPORTED, "libmpdata++: option not supported by the HIP backend");
  if ((p->opts & MPDX_OPT_ABS) && (p->opts & MPDX_OPT_IGA))
    return fail(MPDX_EINVAL, "libmpdata++: abs & iga options are mutually exclusive");
  if (p->n_iters < 1) return fail(MPDX_EINVAL, "libmpdata++: n_iters must be positive");
  if ((p->opts & MPDX_OPT_IGA) && p->n_iters > 2)
    return fail(MPDX_EINVAL, "libmpdata++: infinite gauge option uses just one corrective step");
  if (p->n_eqns < 1 || p->n_eqns > MPDX_MAX_EQNS) return fail(MPDX_EINVAL, "libmpdata++: bad n_eqns");
  for (int d = 0; d < p->n_dims; ++d)
    for (int sd = 0; sd < 2; ++sd)
      if (p->bc[d][sd] != MPDX_BC_CYCLIC && p->bc[d][sd] != MPDX_BC_OPEN)
        return fail(MPDX_EUNSUPPORTED, "libmpdata++: boundary condition not supported by the HIP backend");
  const int hw = (p->opts & MPDX_OPT_FCT) ? 2 : 1;
  for (int d = 0; d < p->n_dims; ++d)
    if (p->grid[d] < hw + 2) return fail(MPDX_EINVAL, "libmpdata++: grid too small for the halo");
  {
    // box kernels put the first index on the grid's y dimension and split
    // the (second, third) plane with 32-bit arithmetic
    const long n1 = p->grid[1] + 2 * hw, n2 = p->n_dims == 3 ? p->grid[2] + 2 * hw : 1;
    if (p->grid[0] + 2 * hw > 65535 || n1 * n2 >= (1l << 31))
      return fail(MPDX_EUNSUPPORTED, "libmpdata++: grid too large for the HIP backend's index space");
  }

  int device = 0;
  {
    std::string err;
    if (int rc = scratch_for_current_device(&device, err)) return fail(rc, err.c_str());
  }
  mpdx_solver *s = new mpdx_solver();
  s->p = *p;
  s->device = device;
  s->esz = p->real_bytes;
  Layout &L = s->L;
  L.nd = p->n_dims;
  L.hw = hw;
  for (int d = 0; d < 3; ++d) L.n[d] = d < L.nd ? p->grid[d] : 1;
  const int align = 128 / s->esz;            // elements per 128 bytes
  const int inner = L.n[L.nd - 1];
  // Offset of index 0 within a row: the boundary kernels touch the strips
  // [-2, 4] and [n-5, n+1] of the contiguous dimension on every row, so
  // place both inside single 128-byte lines where the row length allows it
  // (halves their line traffic); index 0 must stay >= hw from the row start.
  int zoff = align;
  for (int u = 0; u + 7 <= align; ++u) {
    if ((u + inner - 3) % align + 7 <= align) {
      zoff = u + 2;
      break;
    }
  }
  const long pitch = ((long)zoff + inner + hw + align - 1) / align * align;
  if (L.nd == 3) {
    L.st[2] = 1;
    L.st[1] = pitch;
    L.st[0] = pitch * (L.n[1] + 2 * hw);
    L.org = hw * L.st[0] + hw * L.st[1] + zoff;
  } else {
    L.st[2] = 0;
    L.st[1] = 1;
    L.st[0] = pitch;
    L.org = hw * L.st[0] + zoff;
  }
  L.size = L.st[0] * (L.n[0] + 2 * hw);

  hipError_t he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    g_create_error = std::string("libmpdata++: HIP error: ") + hipGetErrorString(he);
    delete s;
    return MPDX_EHIP;
  }
  s->own_stream = true;
  // auxiliary stream (default priority) and fork / join events; without them
  // every launch stays on the main stream
  if (hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    s->use_concurrency = false;
  }
  // run-time switches only turn features off ("0" or empty leaves them on)
  auto off = [](const char *name) {
    const char *ev = std::getenv(name);
    return ev && ev[0] && ev[0] != '0';
  };
  if (off("MPDX_NO_GRAPHS")) s->use_graphs = false;
  if (off("MPDX_NO_CONCURRENCY")) s->use_concurrency = false;
  if (off("MPDX_NO_ZBND")) s->use_zbnd = false;
  if (off("MPDX_NO_SEAM")) s->use_seam = false;
  if (off("MPDX_NO_PASS2D")) s->use_pass2d = false;
  if (off("MPDX_BATCH_EQNS")) s->batch_eqns = true;   // (a switch that turns the batched path on)
  if (off("MPDX_NO_TASK_MIX")) s->mix_tasks = false;
  if (off("MPDX_NO_AUTO_BATCH")) s->auto_batch_eqns = false;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      s->n_cu = ncu;
  }
  if (const char *ev = std::getenv("MPDX_GRAPH_STEPS")) {
    const int g = std::atoi(ev);
    if (g == 1 || (g >= 2 && g <= 64 && g % 2 == 0)) s->graph_steps = g;
  }
  if (const char *ev = std::getenv("MPDX_IX_CHUNK")) {
    const int c = std::atoi(ev);
    if (c >= 8 && c <= 4096) s->ix_chunk = c;
  }

  int rc = 0;
  for (int e = 0; e < p->n_eqns && !rc; ++e) {
    rc = rc || alloc_array(s, &s->psi[e][0]);
    rc = rc || alloc_array(s, &s->psi[e][1]);
  }
  for (int d = 0; d < L.nd && !rc; ++d) rc = alloc_array(s, &s->GC[d]);
  if (!rc) rc = dev_alloc(s, (void **)&s->red_part, sizeof(double2) * (size_t)red_parts(L.n[0]));
  if (!rc) rc = dev_alloc(s, (void **)&s->red_out, sizeof(double));
  if (!rc) {
    const bool fused_ok = fused_supported(s->p, L);
    if (p->path == MPDX_PATH_FUSED && !fused_ok) {
      s->err = "libmpdata++: no fused kernel for this configuration";
      rc = MPDX_EUNSUPPORTED;
    }
    s->active_path = (p->path != MPDX_PATH_PERPASS && fused_ok) ? MPDX_PATH_FUSED : MPDX_PATH_PERPASS;
  }
  if (!rc && s->active_path == MPDX_PATH_FUSED) rc = fused_alloc(s);
  if (!rc) {
    hipError_t e2 = hipStreamSynchronize(s->stream);
    if (e2 != hipSuccess) {
      s->err = std::string("libmpdata++: HIP error: ") + hipGetErrorString(e2);
      rc = MPDX_EHIP;
    }
  }
  if (rc) {
    g_create_error = s->err;
    mpdx_destroy(s);
    return rc;
  }
  *out = s;
  return MPDX_OK;
}

int mpdx_destroy(mpdx_solver *s)
{
  if (!s) return MPDX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_step_graphs(s);
  for (void *p : s->allocs)
    if (p) (void)hipFree(p);
  fused_free(s);
  if (s->aux) {
    (void)hipStreamSynchronize(s->aux);
    (void)hipStreamDestroy(s->aux);
  }
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  if (s->ev_join) (void)hipEventDestroy(s->ev_join);
  for (hipEvent_t ev : s->inflight)
    if (ev) (void)hipEventDestroy(ev);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return MPDX_OK;
}

const char *mpdx_last_error(const mpdx_solver *s)
{
  return s ? s->err.c_str() : g_create_error.c_str();
}

int mpdx_set_stream(mpdx_solver *s, void *stream)
{
  MPDX_NEED(s, "solver handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  // the new stream first, so that a failure leaves the solver on its old one
  hipStream_t next = static_cast<hipStream_t>(stream);
  bool own_next = false;
  if (!next) {
    HIPCHK(hipStreamCreateWithFlags(&next, hipStreamNonBlocking));
    own_next = true;
  }
  hipStream_t old = s->stream;
  const bool own_old = s->own_stream;
  s->stream = next;
  s->own_stream = own_next;
  if (own_old && old) (void)hipStreamDestroy(old);
  return MPDX_OK;
}

int mpdx_advector_extents(const mpdx_solver *s, int d, int ext[3])
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(ext, "output pointer");
  if (d < 0 || d >= s->L.nd) return MPDX_EINVAL;
  int lo[3], hi[3];
  adv_bounds(s, d, lo, hi);
  for (int q = 0; q < 3; ++q) ext[q] = hi[q] - lo[q] + 1;
  return MPDX_OK;
}

int mpdx_set_advectee(mpdx_solver *s, int e, const void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  if (e < 0 || e >= s->p.n_eqns) { s->err = "libmpdata++: bad equation index"; return MPDX_EINVAL; }
  s->seam_known = false;
  int lo[3] = {0, 0, 0}, hi[3] = {s->L.n[0] - 1, s->L.n[1] - 1, s->L.nd == 3 ? s->L.n[2] - 1 : 0};
  return copy_view(s, s->psi[e][s->lev[e]], lo, hi, const_cast<void *>(buf), strides, on_device, true);
}

int mpdx_get_advectee(mpdx_solver *s, int e, void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  if (e < 0 || e >= s->p.n_eqns) { s->err = "libmpdata++: bad equation index"; return MPDX_EINVAL; }
  int lo[3] = {0, 0, 0}, hi[3] = {s->L.n[0] - 1, s->L.n[1] - 1, s->L.nd == 3 ? s->L.n[2] - 1 : 0};
  return copy_view(s, s->psi[e][s->lev[e]], lo, hi, buf, strides, on_device, false);
}

int mpdx_set_advector(mpdx_solver *s, int d, const void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  if (d < 0 || d >= s->L.nd) { s->err = "libmpdata++: bad advector component"; return MPDX_EINVAL; }
  s->seam_known = false;
  int lo[3], hi[3];
  adv_bounds(s, d, lo, hi);
  return copy_view(s, s->GC[d], lo, hi, const_cast<void *>(buf), strides, on_device, true);
}

int mpdx_get_advector(mpdx_solver *s, int d, void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  if (d < 0 || d >= s->L.nd) { s->err = "libmpdata++: bad advector component"; return MPDX_EINVAL; }
  int lo[3], hi[3];
  adv_bounds(s, d, lo, hi);
  return copy_view(s, s->GC[d], lo, hi, buf, strides, on_device, false);
}

int mpdx_advance(mpdx_solver *s, int64_t nt)
{
  MPDX_NEED(s, "solver handle");
  if (nt < 0) { s->err = "libmpdata++: negative number of steps"; return MPDX_EINVAL; }
  if (s->esz == 8) return advance_impl<double, double>(s, nt);
  if (s->p.guard == MPDX_GUARD_NATIVE) return advance_impl<float, float>(s, nt);
  if (s->p.guard == MPDX_GUARD_F64_EXACT) return advance_impl<float, double>(s, nt);
  return advance_impl<float, GuardFast>(s, nt);
}

int mpdx_advance_time(mpdx_solver *s, double duration)
{
  MPDX_NEED(s, "solver handle");
  if (!s->p.var_dt) { s->err = "libmpdata++: advancing by a duration needs var_dt"; return MPDX_EINVAL; }
  if (!(duration >= 0)) { s->err = "libmpdata++: negative duration"; return MPDX_EINVAL; }
  const double t_end = s->time + duration;
  if (s->esz == 8) return advance_impl<double, double>(s, 0, t_end, true);
  if (s->p.guard == MPDX_GUARD_NATIVE) return advance_impl<float, float>(s, 0, t_end, true);
  if (s->p.guard == MPDX_GUARD_F64_EXACT) return advance_impl<float, double>(s, 0, t_end, true);
  return advance_impl<float, GuardFast>(s, 0, t_end, true);
}

int mpdx_dt(const mpdx_solver *s, double *dt)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(dt, "output pointer");
  *dt = s->p.dt;
  return MPDX_OK;
}

int mpdx_reduce(mpdx_solver *s, int e, int op, double *out)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(out, "output pointer");
  if (e < 0 || e >= s->p.n_eqns || op < 0 || op > 2) { s->err = "libmpdata++: bad reduce arguments"; return MPDX_EINVAL; }
  if (s->esz == 8) return reduce_impl<double>(s, s->psi[e][s->lev[e]], op, out);
  return reduce_impl<float>(s, s->psi[e][s->lev[e]], op, out);
}

int mpdx_reduce_noalloc(mpdx_solver *s, int e, int op, double *out)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(out, "output pointer");
  if (e < 0 || e >= s->p.n_eqns || op < 0 || op > 2) { s->err = "libmpdata++: bad reduce arguments"; return MPDX_EINVAL; }
  if (red_parts(s->L.n[0]) > kScratchParts) {
    s->err = "libmpdata++: mpdx_reduce_noalloc: more than " + std::to_string(kScratchParts) +
             " x planes (its fixed scratch holds " + std::to_string(kScratchParts) + " block partials)";
    return MPDX_EINVAL;
  }
  double2 *part = scratch_of(s->device);
  if (!part) { s->err = "libmpdata++: no reduction scratch for this device"; return MPDX_EHIP; }
  if (s->esz == 8) red_field<double>(s->stream, part, out, (const double *)s->psi[e][s->lev[e]], op, s->L);
  else red_field<float>(s->stream, part, out, (const float *)s->psi[e][s->lev[e]], op, s->L);
  HIPCHK(hipGetLastError());
  return MPDX_OK;
}

int mpdx_time(const mpdx_solver *s, double *t)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(t, "output pointer");
  *t = s->time;
  return MPDX_OK;
}

int mpdx_timestep(const mpdx_solver *s, int64_t *ts)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(ts, "output pointer");
  *ts = s->timestep;
  return MPDX_OK;
}

int *mpdx_panic_ptr(mpdx_solver *s)
{
  if (!s) {
    g_create_error = "libmpdata++: null solver handle";
    return nullptr;
  }
  return &s->panic;
}

int mpdx_watch_panic(mpdx_solver *s, const volatile unsigned char *flag)
{
  MPDX_NEED(s, "solver handle");
  s->panic_ext = flag;
  return MPDX_OK;
}

int mpdx_active_path(const mpdx_solver *s) { return s ? s->active_path : -1; }

int mpdx_z_tiles(const mpdx_solver *s)
{
  if (!s) return -1;
  if (s->active_path == MPDX_PATH_FUSED && s->p.n_iters >= 2 && s->use_pass2d && pass2d_supported(s->p, s->L))
    return s->seam == 2 ? 3 : 0;
  if (s->active_path != MPDX_PATH_FUSED || s->p.n_iters < 2 || !interior_supported(s->p, s->L)) return 0;
  const int *b = s->p.bc[2];
  if (s->use_zbnd && b[0] == MPDX_BC_OPEN && b[1] == MPDX_BC_OPEN) return 1;
  if (b[0] != MPDX_BC_CYCLIC || b[1] != MPDX_BC_CYCLIC) return 0;
  return s->seam == 2 ? 3 : (s->seam == 1 ? 2 : 0);
}

int mpdx_set_calc_gc(mpdx_solver *s, mpdx_calc_gc_fn fn, void *user)
{
  MPDX_NEED(s, "solver handle");
  s->seam_known = false;
  s->calc_gc = fn;
  s->calc_gc_user = user;
  return MPDX_OK;
}

int mpdx_set_rhs_scheme(mpdx_solver *s, int scheme, mpdx_update_rhs_fn fn, void *user)
{
  MPDX_NEED(s, "solver handle");
  if (scheme < MPDX_RHS_NONE || scheme > MPDX_RHS_MIXED) {
    s->err = "libmpdata++: unsupported rhs scheme";
    return MPDX_EUNSUPPORTED;
  }
  s->seam_known = false;
  if (scheme != MPDX_RHS_NONE)
    for (int e = 0; e < s->p.n_eqns; ++e)
      if (int rc = alloc_missing(s, &s->rhs[e])) return rc;
  s->rhs_scheme = scheme;
  s->rhs_fn = fn;
  s->rhs_user = user;
  return MPDX_OK;
}

int mpdx_set_rhs_hooks(mpdx_solver *s, mpdx_rhs_hook_fn hook, void *user)
{
  MPDX_NEED(s, "solver handle");
  s->rhs_hook = hook;
  s->rhs_hook_user = user;
  return MPDX_OK;
}

int mpdx_set_var_dt_output(mpdx_solver *s, double outfreq, int ord, int n_vars, const int *vars,
                           mpdx_output_fn fn, void *user)
{
  MPDX_NEED(s, "solver handle");
  if (!fn) {
    s->out_fn = nullptr;
    return MPDX_OK;
  }
  MPDX_NEED(vars, "variable list");
  if (!s->p.var_dt) { s->err = "libmpdata++: interpolated output needs var_dt"; return MPDX_EINVAL; }
  if (!(outfreq > 0) || !std::isfinite(outfreq)) { s->err = "libmpdata++: outfreq must be positive"; return MPDX_EINVAL; }
  if (ord != 1 && ord != 2) {
    s->err = "libmpdata++: only 1st and 2nd order temporal interpolation for output implemented";
    return MPDX_EINVAL;
  }
  if (n_vars < 1 || n_vars > s->p.n_eqns) { s->err = "libmpdata++: bad number of output variables"; return MPDX_EINVAL; }
  for (int v = 0; v < n_vars; ++v) {
    if (vars[v] < 0 || vars[v] >= s->p.n_eqns) { s->err = "libmpdata++: bad equation index"; return MPDX_EINVAL; }
    for (int u = 0; u < v; ++u)
      if (vars[u] == vars[v]) { s->err = "libmpdata++: output variable listed twice"; return MPDX_EINVAL; }
  }
  for (int k = 0; k < ord; ++k)
    for (int v = 0; v < n_vars; ++v)
      if (!s->out_buf[k][v])
        if (int rc = dev_alloc(s, &s->out_buf[k][v], (size_t)s->esz * (size_t)s->L.size)) return rc;
  for (int v = 0; v < n_vars; ++v) s->out_var[v] = vars[v];
  s->out_n = n_vars;
  s->out_ord = ord;
  s->out_freq = outfreq;
  s->out_cnt = 0;
  s->out_fn = fn;
  s->out_user = user;
  return MPDX_OK;
}

int mpdx_get_output(mpdx_solver *s, int e, void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  int v = 0;
  while (v < s->out_n && s->out_var[v] != e) ++v;
  if (!s->out_fn || v == s->out_n) { s->err = "libmpdata++: equation is not an output variable"; return MPDX_EINVAL; }
  int lo[3] = {0, 0, 0}, hi[3] = {s->L.n[0] - 1, s->L.n[1] - 1, s->L.nd == 3 ? s->L.n[2] - 1 : 0};
  return copy_view(s, s->out_buf[0][v], lo, hi, buf, strides, on_device, false);
}

int mpdx_update_rhs(mpdx_solver *s, double dt, int at)
{
  MPDX_NEED(s, "solver handle");
  if (s->rhs_scheme == MPDX_RHS_NONE) { s->err = "libmpdata++: no rhs scheme set"; return MPDX_EINVAL; }
  if (s->esz == 8) return update_rhs<double>(s, dt, at);
  return update_rhs<float>(s, dt, at);
}

int mpdx_apply_rhs(mpdx_solver *s, double dt_arg)
{
  MPDX_NEED(s, "solver handle");
  if (s->rhs_scheme == MPDX_RHS_NONE) { s->err = "libmpdata++: no rhs scheme set"; return MPDX_EINVAL; }
  if (s->esz == 8) apply_rhs<double>(s, dt_arg);
  else apply_rhs<float>(s, dt_arg);
  HIPCHK(hipGetLastError());
  return MPDX_OK;
}

int mpdx_set_rhs(mpdx_solver *s, int e, const void *buf, const int64_t strides[3], int on_device)
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(buf, "buffer");
  MPDX_NEED(strides, "strides");
  if (e < 0 || e >= s->p.n_eqns || !s->rhs[e]) {
    s->err = "libmpdata++: no forcing array for this equation (set an rhs scheme first)";
    return MPDX_EINVAL;
  }
  int lo[3] = {0, 0, 0}, hi[3] = {s->L.n[0] - 1, s->L.n[1] - 1, s->L.nd == 3 ? s->L.n[2] - 1 : 0};
  return copy_view(s, s->rhs[e], lo, hi, const_cast<void *>(buf), strides, on_device, true);
}

static int interior_device(mpdx_solver *s, void *arr, void **ptr, int64_t strides[3])
{
  *ptr = static_cast<char *>(arr) + s->L.org * s->esz;
  for (int q = 0; q < 3; ++q) strides[q] = q < s->L.nd ? s->L.st[q] : 0;
  return MPDX_OK;
}

int mpdx_rhs_device(mpdx_solver *s, int e, void **ptr, int64_t strides[3])
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(ptr, "output pointer");
  MPDX_NEED(strides, "strides");
  if (e < 0 || e >= s->p.n_eqns || !s->rhs[e]) {
    s->err = "libmpdata++: no forcing array for this equation (set an rhs scheme first)";
    return MPDX_EINVAL;
  }
  return interior_device(s, s->rhs[e], ptr, strides);
}

int mpdx_advectee_device(mpdx_solver *s, int e, void **ptr, int64_t strides[3])
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(ptr, "output pointer");
  MPDX_NEED(strides, "strides");
  if (e < 0 || e >= s->p.n_eqns) { s->err = "libmpdata++: bad equation index"; return MPDX_EINVAL; }
  s->seam_exposed = true;
  return interior_device(s, s->psi[e][s->lev[e]], ptr, strides);
}

int mpdx_advector_device(mpdx_solver *s, int d, void **ptr, int64_t strides[3])
{
  MPDX_NEED(s, "solver handle");
  MPDX_NEED(ptr, "output pointer");
  MPDX_NEED(strides, "strides");
  if (d < 0 || d >= s->L.nd) { s->err = "libmpdata++: bad advector component"; return MPDX_EINVAL; }
  s->seam_exposed = true;
  int lo[3], hi[3];
  adv_bounds(s, d, lo, hi);
  const long off = s->L.org + lo[0] * s->L.st[0] + lo[1] * s->L.st[1] + (s->L.nd == 3 ? lo[2] * s->L.st[2] : 0);
  *ptr = static_cast<char *>(s->GC[d]) + off * s->esz;
  for (int q = 0; q < 3; ++q) strides[q] = q < s->L.nd ? s->L.st[q] : 0;
  return MPDX_OK;
}

int mpdx_set_advector_generator(mpdx_solver *s, int kind, const double *params, int n_params)
{
  MPDX_NEED(s, "solver handle");
  // a captured step holds the generator's kernel (or its absence)
  HIPCHK(hipStreamSynchronize(s->stream));
  drop_step_graphs(s);
  s->seam_known = false;
  if (kind == MPDX_GC_NONE) { s->gen_kind = kind; return MPDX_OK; }
  if (s->p.var_dt) {
    s->err = "libmpdata++: the built-in advector generator does not follow a variable dt";
    return MPDX_EUNSUPPORTED;
  }
  if (kind != MPDX_GC_DEFORM || n_params < 2 || !params || !(params[0] > 0) || !(params[1] > 0)) {
    s->err = "libmpdata++: bad advector generator parameters";
    return MPDX_EINVAL;
  }
  // sin^2 tables at corners a + 1/2, a = -2 .. n-1 (host libm, shared with
  // the oracle; a corner is evaluated at its index modulo the period n-1, so
  // the tables repeat bit for bit across the duplicated edge), and the
  // amplitude that makes the largest |GC| (at cos = 1) equal the Courant
  // number
  const int n0 = s->L.n[0], n1 = s->L.n[1];
  const double pi = 3.141592653589793;
  std::vector<double> sx2(n0 + 2), sy2(n1 + 2);
  auto sin2 = [pi](int a, int n) {
    const int r = ((a % (n - 1)) + (n - 1)) % (n - 1);
    const double v = std::sin(pi * (r + 0.5) / (n - 1));
    return v * v;
  };
  for (int a = -2; a < n0; ++a) sx2[a + 2] = sin2(a, n0);
  for (int b = -2; b < n1; ++b) sy2[b + 2] = sin2(b, n1);
  double mx = 0;
  for (int a = -1; a < n0; ++a)
    for (int b = -1; b < n1; ++b) {
      const double c = sx2[a + 2] * sy2[b + 2];
      mx = std::max(mx, std::fabs(c - sx2[a + 2] * sy2[b + 1]));
      mx = std::max(mx, std::fabs(sx2[a + 1] * sy2[b + 2] - c));
    }
  if (!s->gen_sx2) {
    void *px = nullptr, *py = nullptr;
    if (int rc = dev_alloc(s, &px, sizeof(double) * (n0 + 2))) return rc;
    if (int rc = dev_alloc(s, &py, sizeof(double) * (n1 + 2))) return rc;
    s->gen_sx2 = (double *)px;
    s->gen_sy2 = (double *)py;
  }
  HIPCHK(hipMemcpyAsync(s->gen_sx2, sx2.data(), sizeof(double) * (n0 + 2), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->gen_sy2, sy2.data(), sizeof(double) * (n1 + 2), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->gen_kind = kind;
  s->gen_T = params[0];
  s->gen_scale = mx > 0 ? params[1] / mx : 0;
  return MPDX_OK;
}

int mpdx_profile(mpdx_solver *s, int enable)
{
  MPDX_NEED(s, "solver handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  s->prof.clear();
  s->prof_order.clear();
  s->prof_on = enable != 0;
  return MPDX_OK;
}

int mpdx_profile_read(mpdx_solver *s, int idx, char *name, int name_len, double *total_ms,
                      int64_t *launches)
{
  MPDX_NEED(s, "solver handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  if (idx < 0 || idx >= (int)s->prof_order.size()) return MPDX_EINVAL;
  const std::string &n = s->prof_order[idx];
  if (name && name_len > 0) {
    std::strncpy(name, n.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  const mpdx_prof_entry &pe = s->prof[n];
  if (total_ms) *total_ms = pe.ms;
  if (launches) *launches = pe.launches;
  return MPDX_OK;
}

int mpdx_set_x_chunk(mpdx_solver *s, int planes)
{
  MPDX_NEED(s, "solver handle");
  if (planes != 0 && (planes < 8 || planes > 4096)) {
    s->err = "libmpdata++: x chunk must be 0 (automatic) or 8..4096 planes";
    return MPDX_EINVAL;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  s->ix_chunk = planes;
  drop_step_graphs(s);              // captured steps hold the old launch shape
  return MPDX_OK;
}

int mpdx_set_tile_y(mpdx_solver *s, int rows)
{
  MPDX_NEED(s, "solver handle");
  if (rows != 0 && rows != 16 && rows != 32) {
    s->err = "libmpdata++: interior tile height must be 0 (default), 16 or 32";
    return MPDX_EINVAL;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  s->tile_y = rows;
  drop_step_graphs(s);              // captured steps hold the old kernels
  return MPDX_OK;
}

int mpdx_tune(mpdx_solver *s, int steps_per_trial, int64_t *steps_taken, int *tile_y, int *x_chunk)
{
  MPDX_NEED(s, "solver handle");
  if (steps_taken) *steps_taken = 0;
  if (steps_per_trial < 1) { s->err = "libmpdata++: steps_per_trial must be >= 1"; return MPDX_EINVAL; }
  auto report = [&]() {
    if (tile_y) *tile_y = s->tile_y;
    if (x_chunk) *x_chunk = s->ix_chunk;
  };
  // only the 3D interior kernel has these knobs
  if (s->active_path != MPDX_PATH_FUSED || s->L.nd != 3 || !interior_supported(s->p, s->L)) {
    report();
    return MPDX_OK;
  }
  int span = 16;
  while (span < s->L.n[0] - 4) span *= 2;
  const bool tiles = s->esz == 4 && s->p.guard == MPDX_GUARD_F64;
  std::vector<std::pair<int, int>> trials;   // (tile height, x chunk)
  for (int ty : tiles ? std::vector<int>{0, 32, 16} : std::vector<int>{0})
    for (int c : {0, 16, 32, 64, 128, 256})
      if (c <= span) trials.emplace_back(ty, c);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  HIPCHK(hipEventCreate(&t0));
  if (hipEventCreate(&t1) != hipSuccess) {
    (void)hipEventDestroy(t0);
    s->err = "libmpdata++: cannot create timing events";
    return MPDX_EHIP;
  }
  const std::pair<int, int> start(s->tile_y, s->ix_chunk);
  std::pair<int, int> best = start;
  float best_ms = -1;
  int64_t taken = 0;
  int rc = 0;
  for (const auto &tr : trials) {
    if ((rc = mpdx_set_tile_y(s, tr.first)) || (rc = mpdx_set_x_chunk(s, tr.second))) break;
    if ((rc = mpdx_advance(s, 2))) break;   // captures the step graphs
    taken += 2;
    float ms = 0;
    if (hipEventRecord(t0, s->stream) != hipSuccess) rc = MPDX_EHIP;
    if (!rc) rc = mpdx_advance(s, steps_per_trial);
    if (!rc && (hipEventRecord(t1, s->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess
                || hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
      rc = MPDX_EHIP;
    if (rc) break;
    taken += steps_per_trial;
    if (best_ms < 0 || ms < best_ms) {
      best_ms = ms;
      best = tr;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (steps_taken) *steps_taken = taken;
  if (rc == MPDX_EHIP && s->err.empty()) s->err = "libmpdata++: HIP error while tuning";
  // keep the fastest pair (the starting one if no trial finished)
  if (best_ms < 0) best = start;
  const std::string err = s->err;
  int rc2 = mpdx_set_tile_y(s, best.first);
  if (!rc2) rc2 = mpdx_set_x_chunk(s, best.second);
  if (rc) s->err = err;
  report();
  return rc ? rc : rc2;
}

int mpdx_ensemble_create(mpdx_solver *const *members, int n, mpdx_ensemble **out)
{
  MPDX_NEED(out, "output pointer");
  *out = nullptr;
  auto reject = [&](const std::string &msg) {
    g_create_error = "libmpdata++: " + msg;
    return MPDX_EINVAL;
  };
  if (n < 1 || !members) return reject("an ensemble needs at least one member");
  for (int i = 0; i < n; ++i) {
    const mpdx_solver *s = members[i];
    const std::string who = "ensemble member " + std::to_string(i);
    if (!s) return reject(who + " is null");
    for (int j = 0; j < i; ++j)
      if (members[j] == s) return reject(who + " appears twice");
    if (s->active_path != MPDX_PATH_FUSED) return reject(who + " does not use the fused path");
    if (s->calc_gc || s->rhs_scheme != MPDX_RHS_NONE || s->gen_kind != MPDX_GC_NONE)
      return reject(who + " has a calc_gc hook, forcing or an advector generator");
    if (s->p.upwind_filter_freq > 0 || s->p.var_dt) return reject(who + " uses upwind_filter_freq or var_dt");
    if (!s->stream) return reject(who + " runs on the legacy null stream");
    if (s->device != members[0]->device) return reject(who + " lives on another device");
  }
  mpdx_ensemble *e = new mpdx_ensemble;
  e->m.assign(members, members + n);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    e->stream = nullptr;
    mpdx_ensemble_destroy(e);
    g_create_error = "libmpdata++: cannot create the ensemble's stream";
    return MPDX_EHIP;
  }
  int ns = 1;
  if (const char *v = std::getenv("MPDX_ENS_STREAMS"))
    if (atoi(v) > 0) ns = std::min(atoi(v), mpdx_ensemble::kAux + 1);
  for (int k = 0; k + 1 < ns; ++k) {
    if (hipStreamCreateWithFlags(&e->aux[k], hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      e->aux[k] = nullptr;
      break;   // fewer streams: the batches still run, with less overlap
    }
    e->n_aux = k + 1;
  }
  *out = e;
  return MPDX_OK;
}

int mpdx_ensemble_destroy(mpdx_ensemble *e)
{
  if (!e) return MPDX_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_plans(e);
  for (hipEvent_t ev : e->inflight)
    if (ev) (void)hipEventDestroy(ev);
  for (hipStream_t a : e->aux)
    if (a) (void)hipStreamDestroy(a);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return MPDX_OK;
}

const char *mpdx_ensemble_last_error(const mpdx_ensemble *e)
{
  return e ? e->err.c_str() : g_create_error.c_str();
}

int mpdx_ensemble_advance(mpdx_ensemble *e, int64_t nt)
{
  MPDX_NEED(e, "ensemble handle");
  if (nt < 0) { e->err = "libmpdata++: negative number of steps"; return MPDX_EINVAL; }
  if (nt == 0) return MPDX_OK;
  if (int rc = ensemble_ante(e)) return rc;
  if (int rc = ensemble_run(e, nt)) return rc;
  EHIPCHK(hipStreamSynchronize(e->stream));
  return MPDX_OK;
}

int mpdx_ensemble_stats(const mpdx_ensemble *e, int *launches, int64_t *tasks)
{
  MPDX_NEED(e, "ensemble handle");
  const EnsPlan &p = e->plan[0].graph ? e->plan[0] : e->plan[1];
  if (launches) *launches = (int)p.bp.batches.size();
  if (tasks) *tasks = (int64_t)p.bp.tasks;
  return MPDX_OK;
}

int mpdx_ensemble_tune(mpdx_ensemble *e, int steps_per_trial, int64_t *steps_taken)
{
  MPDX_NEED(e, "ensemble handle");
  if (steps_taken) *steps_taken = 0;
  if (steps_per_trial < 1) { e->err = "libmpdata++: steps_per_trial must be >= 1"; return MPDX_EINVAL; }
  if (int rc = ensemble_ante(e)) return rc;
  // classes: 3D members by x extent rounded up to a power of two; the FP32
  // default-guard ones may take either interior tile height
  std::map<int, std::vector<mpdx_solver *>> cls;
  for (mpdx_solver *s : e->m)
    if (s->L.nd == 3 && interior_supported(s->p, s->L)) {
      int b = 16;
      while (b < s->L.n[0] - 4) b *= 2;
      cls[b].push_back(s);
    }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  EHIPCHK(hipEventCreate(&t0));
  if (hipEventCreate(&t1) != hipSuccess) {
    (void)hipEventDestroy(t0);
    e->err = "libmpdata++: cannot create timing events";
    return MPDX_EHIP;
  }
  int rc = 0;
  int64_t taken = 0;
  for (auto &kv : cls) {
    bool tiles = true;
    for (mpdx_solver *s : kv.second) tiles = tiles && s->esz == 4 && s->p.guard == MPDX_GUARD_F64;
    std::vector<std::pair<int, int>> trials;   // (tile height, x chunk)
    for (int ty : tiles ? std::vector<int>{0, 16} : std::vector<int>{0})
      for (int c : {0, 16, 32, 64, 128, 256})
        if (c <= kv.first) trials.emplace_back(ty, c);
    std::pair<int, int> best(kv.second.front()->tile_y, kv.second.front()->ix_chunk);
    float best_ms = -1;
    for (const auto &tr : trials) {
      for (mpdx_solver *s : kv.second) {
        s->tile_y = tr.first;
        s->ix_chunk = tr.second;
      }
      if (hipStreamSynchronize(e->stream) != hipSuccess) rc = MPDX_EHIP;
      drop_plans(e);
      if (!rc) rc = ensemble_run(e, 2);   // builds both level states' plans
      if (!rc && hipEventRecord(t0, e->stream) != hipSuccess) rc = MPDX_EHIP;
      if (!rc) rc = ensemble_run(e, steps_per_trial);
      if (!rc && hipEventRecord(t1, e->stream) != hipSuccess) rc = MPDX_EHIP;
      float ms = 0;
      if (!rc && (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
        rc = MPDX_EHIP;
      if (rc) break;
      taken += 2 + steps_per_trial;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        best = tr;
      }
    }
    for (mpdx_solver *s : kv.second) {
      s->tile_y = best.first;
      s->ix_chunk = best.second;
      drop_step_graphs(s);
    }
    if (rc) break;
  }
  // classes that chose different tile heights run their interior passes as
  // different batched launches (a mixed-variant task holds one tile
  // height): time the chosen set against one height for every class that
  // can take it (keeping the chunks), and keep the fastest
  auto trial = [&](float *ms) {
    int r = 0;
    if (hipStreamSynchronize(e->stream) != hipSuccess) r = MPDX_EHIP;
    drop_plans(e);
    if (!r) r = ensemble_run(e, 2);
    if (!r && hipEventRecord(t0, e->stream) != hipSuccess) r = MPDX_EHIP;
    if (!r) r = ensemble_run(e, steps_per_trial);
    if (!r && hipEventRecord(t1, e->stream) != hipSuccess) r = MPDX_EHIP;
    if (!r && (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(ms, t0, t1) != hipSuccess)) r = MPDX_EHIP;
    if (!r) taken += 2 + steps_per_trial;
    return r;
  };
  std::vector<mpdx_solver *> tunable;
  for (auto &kv : cls)
    for (mpdx_solver *s : kv.second)
      if (s->esz == 4 && s->p.guard == MPDX_GUARD_F64) tunable.push_back(s);
  bool mixed = false;
  for (mpdx_solver *s : tunable) mixed = mixed || s->tile_y != tunable.front()->tile_y;
  if (!rc && mixed) {
    std::vector<int> chosen;
    for (mpdx_solver *s : tunable) chosen.push_back(s->tile_y);
    float best_ms = 0;
    int best_ty = -1;   // -1: the per-class choice
    rc = trial(&best_ms);
    for (int ty : {0, 16}) {
      if (rc) break;
      for (mpdx_solver *s : tunable) s->tile_y = ty;
      float ms = 0;
      rc = trial(&ms);
      if (!rc && ms < best_ms) {
        best_ms = ms;
        best_ty = ty;
      }
    }
    for (size_t i = 0; i < tunable.size(); ++i) {
      tunable[i]->tile_y = best_ty < 0 ? chosen[i] : best_ty;
      drop_step_graphs(tunable[i]);
    }
  }
  (void)hipStreamSynchronize(e->stream);
  drop_plans(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (steps_taken) *steps_taken = taken;
  if (rc == MPDX_EHIP && e->err.empty()) e->err = "libmpdata++: HIP error while tuning";
  return rc;
}

int mpdx_ensemble_x_chunk(const mpdx_ensemble *e, int i, int *planes)
{
  MPDX_NEED(e, "ensemble handle");
  MPDX_NEED(planes, "output pointer");
  if (i < 0 || i >= (int)e->m.size()) return MPDX_EINVAL;
  *planes = e->m[i]->ix_chunk;
  return MPDX_OK;
}

int mpdx_ensemble_tile_y(const mpdx_ensemble *e, int i, int *rows)
{
  MPDX_NEED(e, "ensemble handle");
  MPDX_NEED(rows, "output pointer");
  if (i < 0 || i >= (int)e->m.size()) return MPDX_EINVAL;
  const mpdx_solver *s = e->m[i];
  // FP64: 24 rows; a single corrective pass (n_iters = 2) 16; FP32 32 (an
  // ensemble's batched launches keep 32 rows for every pass)
  *rows = s->tile_y ? s->tile_y
                    : (s->esz == 8 ? (s->p.n_iters == 2 ? itile_y_default<double, true, false>()
                                                        : itile_y_default<double>())
                                   : itile_y_default<float>());
  return MPDX_OK;
}

}  // extern "C"

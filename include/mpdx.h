/*
 * mpdx.h -- C ABI of the MI355X MPDATA advection backend.
 *
 * This is the drop-in boundary behind libmpdata++'s concurrency-backend slot:
 * the header-only C++ template libmpdataxx::concurr::hip<solver_t, bcond...>
 * (include/libmpdata++/concurr/hip.hpp) calls only these functions, exactly
 * where the reference's concurr::serial/openmp would run the solver objects.
 * Plain pointers and sizes only; all calls return 0 on success and a nonzero
 * MPDX_E* code on failure, with the message (libmpdata++'s own wording where
 * the reference throws) available from mpdx_last_error().  A null handle or
 * output pointer gives MPDX_EINVAL, its message from mpdx_last_error(NULL)
 * (mpdx_ensemble_last_error(NULL) for ensembles).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   mpdx_create            concurr::serial ctor -> concurr_common ctor
 *                          (concurr/serial.hpp:49-51,
 *                           concurr/detail/concurr_common.hpp:176-187),
 *                          solver_common ctor checks (solvers/detail/
 *                          solver_common.hpp:246-248)
 *   mpdx_advance           concurr_common::advance -> solver_common::solve
 *                          (concurr_common.hpp:380-385, solver_common.hpp:270-358)
 *   mpdx_set/get_advectee  advectee(e) / advectee_global(_set)
 *                          (concurr_common.hpp:387-408, sharedmem.hpp:411-419,522-531)
 *   mpdx_set/get_advector  advector(d) (concurr_common.hpp:410-413,
 *                          sharedmem.hpp:431-456,543-577)
 *   mpdx_reduce            min(e)/max(e) (concurr_common.hpp:445-453) and the
 *                          double-precision Kahan sum of sharedmem::sum
 *                          (sharedmem.hpp:115-155, kahan_reduction.hpp:15-58)
 *   mpdx_time / mpdx_timestep   time() (concurr_common.hpp:440-443)
 *   mpdx_panic_ptr         panic_ptr() (concurr_common.hpp:435-438)
 */
#ifndef MPDX_H
#define MPDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPDX_MAX_EQNS 64

/* option bits: libmpdata++/opts.hpp:34-49 (same values) */
#define MPDX_OPT_FCT (1u << 0)
#define MPDX_OPT_ABS (1u << 1)
#define MPDX_OPT_TOT (1u << 2)
#define MPDX_OPT_PFC (1u << 3)
#define MPDX_OPT_NPA (1u << 4)
#define MPDX_OPT_IGA (1u << 5)
#define MPDX_OPT_NUG (1u << 6)
#define MPDX_OPT_DFL (1u << 7)
#define MPDX_OPT_KHN (1u << 8)

/* boundary kinds: bcond::cyclic / bcond::open (bcond/detail/bcond_common.hpp:17) */
#define MPDX_BC_CYCLIC 0
#define MPDX_BC_OPEN 1

/* precision of state, advector and fluxes */
#define MPDX_REAL_F32 4
#define MPDX_REAL_F64 8

/* How the near-zero-denominator guards of the psi ratios (n/(d + tiny)) and
 * FCT betas (n/(d + epsilon)) are evaluated on the FP32 path
 * (formulae/mpdata/formulae_mpdata_common.hpp:61-98).  FP64 solvers always use
 * the reference's FP64 semantics.
 *   MPDX_GUARD_F64        FP64 guard constants, ratio = n * rcp(d + guard) in
 *                         FP32 with the hardware reciprocal (default)
 *   MPDX_GUARD_NATIVE     numeric_limits<float> constants and IEEE FP32
 *                         division: libmpdata++ with real_t = float
 *   MPDX_GUARD_F64_EXACT  operands promoted to FP64, FP64 constants, IEEE FP64
 *                         division, result rounded to FP32 */
#define MPDX_GUARD_F64 0
#define MPDX_GUARD_NATIVE 1
#define MPDX_GUARD_F64_EXACT 2

/* execution path */
#define MPDX_PATH_AUTO 0      /* fused kernels where available, else per-pass */
#define MPDX_PATH_PERPASS 1   /* one kernel per reference pass, literal halo fills */
#define MPDX_PATH_FUSED 2     /* fused stencil kernels (error if unsupported) */

/* error codes */
#define MPDX_OK 0
#define MPDX_EINVAL 1
#define MPDX_EDIVERGENT 2
#define MPDX_EHIP 3
#define MPDX_ENOMEM 4
#define MPDX_EUNSUPPORTED 5

/* reductions */
#define MPDX_RED_MIN 0
#define MPDX_RED_MAX 1
#define MPDX_RED_SUM 2        /* FP64 Kahan-compensated sum over the interior */

typedef struct mpdx_solver mpdx_solver;

typedef struct mpdx_params {
  int n_dims;                     /* 2 or 3 (ct_params_t::n_dims) */
  int grid[3];                    /* rt_params_t::grid_size */
  int real_bytes;                 /* MPDX_REAL_F32 or MPDX_REAL_F64 (ct_params_t::real_t) */
  uint32_t opts;                  /* ct_params_t::opts */
  int n_iters;                    /* rt_params_t::n_iters (mpdata_common.hpp:74-78) */
  int n_eqns;                     /* ct_params_t::n_eqns: tracers sharing one advector */
  int bc[3][2];                   /* [dim][left,right] MPDX_BC_* */
  int upwind_filter_freq;         /* rt_params_t::upwind_filter_freq */
  int hint_scale[MPDX_MAX_EQNS];  /* ct_params_t::hint_scale(e) */
  double dt;                      /* rt_params_t::dt */
  double max_abs_div_eps;         /* rt_params_t::max_abs_div_eps; < 0 = epsilon<real_t> */
  int guard;                      /* MPDX_GUARD_* */
  int path;                       /* MPDX_PATH_* */
  int var_dt;                     /* ct_params_t::var_dt: adaptive time step */
  double max_courant;             /* rt_params_t::max_courant (solver_common.hpp:216-220) */
} mpdx_params;

/* fills p with the reference defaults (opts = iga|fct, n_iters = 2, ...) */
void mpdx_params_default(mpdx_params *p);

int mpdx_create(const mpdx_params *p, mpdx_solver **out);
int mpdx_destroy(mpdx_solver *h);
/* message of the last failed call on h (h may be NULL for mpdx_create) */
const char *mpdx_last_error(const mpdx_solver *h);

/* hipStream_t to enqueue on (NULL = the library's own stream) */
int mpdx_set_stream(mpdx_solver *h, void *stream);

/* Advectee/advector copies.  buf holds real_bytes-wide elements; strides are
 * in elements, for [x][y][z]; on_device != 0 means buf is device memory.
 * The advectee covers the interior (grid); the advector(d) view covers the
 * same indices as the reference's advector(d): faces -1..n_d-2 along d and,
 * for 3D components d > 0, x = -1..n0-1 (buf points at that first element).
 * With on_device the call first waits for all work enqueued on the device
 * (any stream, so a buffer still being produced elsewhere is complete) and
 * returns when the copy is done; host copies return when done as well. */
int mpdx_advector_extents(const mpdx_solver *h, int d, int ext[3]);
int mpdx_set_advectee(mpdx_solver *h, int e, const void *buf, const int64_t strides[3], int on_device);
int mpdx_get_advectee(mpdx_solver *h, int e, void *buf, const int64_t strides[3], int on_device);
int mpdx_set_advector(mpdx_solver *h, int d, const void *buf, const int64_t strides[3], int on_device);
int mpdx_get_advector(mpdx_solver *h, int d, void *buf, const int64_t strides[3], int on_device);

/* advance by nt time steps (synchronous: returns when the work is done).
 * A var_dt solver takes nt steps of its current dt (with the same
 * Courant-number adjustment as mpdx_advance_time). */
int mpdx_advance(mpdx_solver *h, int64_t nt);

/* var_dt: advance by a duration (advance(real_t) of a var_dt solver,
 * solver_common.hpp:270-274,292-355).  On the first call dt and the
 * advector are rescaled so that the Courant number (solver_3d.hpp:390-398)
 * equals max_courant (hook_ante_loop, solver_common.hpp:199-210); with a
 * calc_gc hook dt is re-adjusted every step until the Courant number is at
 * most max_courant.  Steps are taken while time < time at the call +
 * duration, and at least one: the reference's loop runs out_intrp_ord
 * (default 1) further steps once time has reached the end, so
 * advance(0) takes one step and a step that ends at or past the end is
 * the last.  The C++ wrapper takes the second extra step of
 * out_intrp_ord == 2 with mpdx_advance(h, 1).  Refused (EINVAL) for
 * constant-dt solvers.  mpdx_dt returns the current dt. */
int mpdx_advance_time(mpdx_solver *h, double duration);
int mpdx_dt(const mpdx_solver *h, double *dt);

int mpdx_reduce(mpdx_solver *h, int e, int op, double *out);
/* No counterpart in the reference: mpdx_reduce for callers that cannot
 * allocate or block (code under HIP graph capture, tight serving loops).
 * Same kernels and result bit for bit, but the block partials live in a
 * fixed 128 KiB device array per GPU (nothing is allocated), and the result
 * is written by the GPU to out (device memory, or host memory the device can
 * write) in the solver's stream order: the call returns without waiting.
 * Grids with more than 8192 x planes are refused (MPDX_EINVAL).  Every call
 * on one GPU uses the same array: calls that overlap in time on one GPU
 * (different streams, threads or solvers) race and must be ordered by the
 * caller. */
int mpdx_reduce_noalloc(mpdx_solver *h, int e, int op, double *out);
int mpdx_time(const mpdx_solver *h, double *t);
int mpdx_timestep(const mpdx_solver *h, int64_t *ts);
/* host flag polled once per step; set nonzero to make advance() return early.
 * advance() keeps at most a few steps queued on the GPU ahead of the host, so
 * a flag set while it runs (another thread, a signal handler) stops it within
 * those steps; the steps already queued complete. */
int *mpdx_panic_ptr(mpdx_solver *h);
/* a second flag, owned by the caller (a byte, nonzero = stop), polled like
 * the library's own: the C++ backend registers its bool panic_ so that the
 * reference's panic_ptr() (concurr_common.hpp:435-438, polled by
 * solver_common::solve, solver_common.hpp:301-302) works across one
 * multi-step advance() call.  NULL removes it.  The flag must outlive the
 * solver or be removed first. */
int mpdx_watch_panic(mpdx_solver *h, const volatile unsigned char *flag);
/* which path advance() uses for this configuration (MPDX_PATH_PERPASS/FUSED;
 * -1 for a null handle) */
int mpdx_active_path(const mpdx_solver *h);
/* how the last advance() call's interior kernel treated the z bounds (3D
 * fused path; no counterpart in the reference, which has no such kernel):
 * 0 = the boundary shell computed the cells next to them, 1 = open bounds on
 * both z sides, computed in the interior kernel's tiles, 2 = cyclic bounds on
 * both z sides whose fields repeat plane z = 0 at z = n-1, read by the tiles
 * through the periodic images of the halo columns, 3 = cyclic bounds in all
 * three dimensions with fields repeating plane 0 at n-1 along each: the tiles
 * compute every cell and there is no boundary shell (both decided at the
 * start of each call by comparing those planes on the device); 2D: 0, or 3
 * when the 2D pass kernel's tiles cover every cell of such an all-cyclic
 * domain; -1 for a null handle */
int mpdx_z_tiles(const mpdx_solver *h);

/* Time-varying advector (the reference's calc_gc() hook, solver_common.hpp:
 * 119-120,308, e.g. tests/mp3_paper_2018_JCP/reversing_deform/
 * reversing_deform.hpp:64-105,172-179).  fn is called once per time step,
 * before the first equation is advected, with the mid-step time t + dt/2; it
 * may write the advector with mpdx_set_advector or through the pointers of
 * mpdx_advector_device (kernels on the solver's stream).  A nonzero return
 * means the advector changed: the library then refills its halos exactly as
 * the reference's calc_gc implementations do (xchng_vctr_alng followed by
 * xchng_vctr_nrml with ext = halo - 1).  fn = NULL removes the hook. */
typedef int (*mpdx_calc_gc_fn)(mpdx_solver *h, double t_half, void *user);
int mpdx_set_calc_gc(mpdx_solver *h, mpdx_calc_gc_fn fn, void *user);

/* device address of the first element of the advector(d) view and its
 * element strides (same indexing as mpdx_set_advector) */
int mpdx_advector_device(mpdx_solver *h, int d, void **ptr, int64_t strides[3]);

/* Built-in time-varying advector computed on the GPU at every step (no
 * reference counterpart; it reproduces the calc_gc pattern above without a
 * host round trip).  MPDX_GC_DEFORM: the reversing deformational flow of the
 * streamfunction sin^2(pi x) sin^2(pi y) cos(pi t / T) / pi with
 * x = (i + 1/2)/(n0 - 1), y = (j + 1/2)/(n1 - 1) at cell corners (period n-1,
 * consistent with cyclic boundaries), GC = scaled discrete curl, z-independent
 * in 3D (GC_z = 0).  params = {T, max_courant}; corner values are rounded to
 * multiples of 2^-17 so the discrete divergence is exactly zero. */
#define MPDX_GC_NONE 0
#define MPDX_GC_DEFORM 1
int mpdx_set_advector_generator(mpdx_solver *h, int kind, const double *params, int n_params);

/* Explicit forcing: the rhs schemes of solvers::mpdata_rhs
 * (solvers/mpdata_rhs.hpp:20-25,116-195): euler_a psi^{n+1} = ADV(psi^n) +
 * dt R^n; euler_b psi^{n+1} = ADV(psi^n + dt R^n); trapez psi^{n+1} =
 * ADV(psi^n + dt/2 R^n) + dt/2 R^{n+1}.  fn plays the user's update_rhs(rhs,
 * dt, at) (at = 0: R^n, at = 1: R^{n+1}, trapez only): before each call the
 * forcing arrays are zeroed over the interior, as the reference's base
 * update_rhs does (mpdata_rhs.hpp:52-70); fn fills them with mpdx_set_rhs or
 * through mpdx_rhs_device (kernels on the solver's stream), reading the state
 * with mpdx_get_advectee or mpdx_advectee_device.  Forcing is added over the interior (apply_rhs,
 * mpdata_rhs.hpp:72-84).  hint_norhs is not supported. */
#define MPDX_RHS_NONE 0
#define MPDX_RHS_EULER_A 1
#define MPDX_RHS_EULER_B 2
#define MPDX_RHS_TRAPEZ 3
#define MPDX_RHS_MIXED 4
typedef int (*mpdx_update_rhs_fn)(mpdx_solver *h, double dt, int at, void *user);
int mpdx_set_rhs_scheme(mpdx_solver *h, int scheme, mpdx_update_rhs_fn fn, void *user);
/* The mixed scheme (rhs_scheme_t::mixed, mpdata_rhs.hpp:24,112-124): the
 * library applies no forcing itself but calls hook(h, MPDX_HOOK_*, user) where
 * the reference calls hook_mixed_rhs_ante_loop / _ante_step / _post_step
 * (the first advance's hook_ante_loop, before and after every step), and the
 * hook implements the scheme with the two calls below: mpdx_update_rhs (the
 * reference's update_rhs: zero the forcing over the interior, then the
 * scheme's update function fills it) and mpdx_apply_rhs (apply_rhs: state
 * += dt_arg * forcing over the interior).  Both also work for the other
 * schemes. */
#define MPDX_HOOK_ANTE_LOOP 0
#define MPDX_HOOK_ANTE_STEP 1
#define MPDX_HOOK_POST_STEP 2
typedef int (*mpdx_rhs_hook_fn)(mpdx_solver *h, int hook, void *user);
int mpdx_set_rhs_hooks(mpdx_solver *h, mpdx_rhs_hook_fn hook, void *user);
int mpdx_update_rhs(mpdx_solver *h, double dt, int at);
int mpdx_apply_rhs(mpdx_solver *h, double dt_arg);
/* Output with var_dt (output/detail/output_common.hpp:32-37,47-62,72-140):
 * records at the times k * outfreq, interpolated in time from the states
 * around them (intrp_ord 1: linear from the states before and after the
 * step that reaches k * outfreq; 2: quadratic, with the state after one
 * more step).  The library keeps the interpolation buffers on the device
 * for the equations listed in vars and calls fn(h, kind, t, user):
 *   MPDX_OUT_START:  hook_ante_loop of the first advance from timestep 0,
 *                    after dt was rescaled; t = the end time of that call
 *                    (time + duration; time + nt * dt for mpdx_advance);
 *                    the buffers hold the current state;
 *   MPDX_OUT_RECORD: after the step that completes a record (hook_post_step),
 *                    with the solver's timestep and time already advanced;
 *                    t = the record's time k * outfreq.
 * In the callback mpdx_get_output copies the buffer of equation e (interior,
 * like mpdx_get_advectee).  A nonzero return from fn stops the advance with
 * MPDX_EINVAL.  Only for var_dt solvers; fn = NULL removes the schedule.
 * Times are compared and coefficients formed in double. */
#define MPDX_OUT_START 0
#define MPDX_OUT_RECORD 1
typedef int (*mpdx_output_fn)(mpdx_solver *h, int kind, double t, void *user);
int mpdx_set_var_dt_output(mpdx_solver *h, double outfreq, int intrp_ord, int n_vars, const int *vars,
                           mpdx_output_fn fn, void *user);
int mpdx_get_output(mpdx_solver *h, int e, void *buf, const int64_t strides[3], int on_device);

/* copy the forcing of equation e (interior, like mpdx_set_advectee) */
int mpdx_set_rhs(mpdx_solver *h, int e, const void *buf, const int64_t strides[3], int on_device);
/* device address of element (0, 0, 0) of the forcing / the current state of
 * equation e, and the element strides */
int mpdx_rhs_device(mpdx_solver *h, int e, void **ptr, int64_t strides[3]);
int mpdx_advectee_device(mpdx_solver *h, int e, void **ptr, int64_t strides[3]);

/* Kernel timing with HIP events on the solver's stream (no reference
 * counterpart; the reference prints a wall-clock timer, concurr/detail/
 * timer.hpp:19-56).  mpdx_profile(h, 1) clears and starts recording;
 * mpdx_profile_read returns, for the idx-th kernel class recorded, its name,
 * summed device time and launch count (MPDX_EINVAL past the last class). */
int mpdx_profile(mpdx_solver *h, int enable);
int mpdx_profile_read(mpdx_solver *h, int idx, char *name, int name_len,
                      double *total_ms, int64_t *launches);

/* Length, in output planes, of the x chunks the 3D interior kernel marches
 * through (0 = automatic: as long as possible while leaving two tiles per
 * CU).  A tuning knob of this backend with no reference counterpart; results
 * do not depend on it. */
int mpdx_set_x_chunk(mpdx_solver *h, int planes);

/* Interior tile height of the 3D interior kernel for FP32 solvers with the
 * default guards: 32 (one 1024-thread block per CU) or 16 rows (two
 * 512-thread blocks per CU) for every corrective pass; 0 = the default: 32
 * rows for the first corrective pass, 24 (two 768-thread blocks per CU) for
 * the later ones (32 for all of them when the solver is stepped as an
 * ensemble member).  FP64 and the other guard modes always use their own
 * default.  A tuning knob with no reference counterpart; results do not
 * depend on it. */
int mpdx_set_tile_y(mpdx_solver *h, int rows);

/* Autotuner for one solver (no reference counterpart): tries the interior
 * tile heights (FP32 with the default guards: 0, 32 and 16 rows; otherwise
 * the default) with the x-chunk lengths 0, 16, 32, 64, 128, 256 (those up to
 * the x extent rounded up to a power of two), timing steps_per_trial steps
 * of each after two untimed ones (which capture the step graphs), and keeps
 * the fastest pair.  The trial steps are real steps of the solver (use them
 * as warm-up); *steps_taken returns their number, *tile_y and *x_chunk the
 * pair kept (each output may be NULL).  A solver without the 3D interior
 * kernel (2D, the per-pass path, options the kernel does not cover) takes
 * no steps.  Results do not depend on the pair. */
int mpdx_tune(mpdx_solver *h, int steps_per_trial, int64_t *steps_taken, int *tile_y, int *x_chunk);

/* Ensembles: independent solvers (domains of any size, dimensionality,
 * options and precision) advanced together, as the reference advances each
 * solver object with its own advance(nt) (concurr_common.hpp:380-385).  One
 * time step of every member is recorded as the launches its own step makes
 * (same kernels, same arguments), launches of the same kernel are merged
 * across members into single batched launches (a device table of per-member
 * tasks), and the batched launches are captured on the ensemble's own stream
 * as one linear hipGraph that is replayed once per step (MPDX_ENS_STREAMS=n
 * spreads them over n streams by member dependencies instead).  Members stay owned
 * by the caller, must outlive the ensemble, live on one device, and keep
 * working on their own between ensemble calls (their pending work is
 * finished before an ensemble step starts).  They must use the fused path
 * and have no host callback (calc_gc, forcing), no advector generator, no
 * upwind filter and no var_dt.  The panic flags are polled between steps. */
typedef struct mpdx_ensemble mpdx_ensemble;
int mpdx_ensemble_create(mpdx_solver *const *members, int n, mpdx_ensemble **out);
int mpdx_ensemble_destroy(mpdx_ensemble *e);
int mpdx_ensemble_advance(mpdx_ensemble *e, int64_t nt);
/* Batched launches and tasks of one ensemble step (of the last step's plan) */
int mpdx_ensemble_stats(const mpdx_ensemble *e, int *launches, int64_t *tasks);
/* Autotuner (no reference counterpart): for each class of 3D members (x
 * extent rounded up to a power of two) tries the interior tile heights
 * (FP32 default-guard classes: 32 and 16 rows) with the x-chunk lengths 0,
 * 16, 32, 64, 128, 256 (those that fit), timing steps_per_trial ensemble
 * steps of each after two untimed ones, and keeps the fastest pair; if the
 * classes chose different heights, it then times one height (the default,
 * then 16 rows) for all of them against that choice and keeps the fastest
 * (one height lets every member's interior pass share a batched launch).
 * The trial steps are real steps of every member (use them as warm-up);
 * *steps_taken returns their number. */
int mpdx_ensemble_tune(mpdx_ensemble *e, int steps_per_trial, int64_t *steps_taken);
/* x-chunk length and interior tile height chosen for member i */
int mpdx_ensemble_x_chunk(const mpdx_ensemble *e, int i, int *planes);
int mpdx_ensemble_tile_y(const mpdx_ensemble *e, int i, int *rows);
const char *mpdx_ensemble_last_error(const mpdx_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif /* MPDX_H */

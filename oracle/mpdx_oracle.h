/*
 * mpdx_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of libmpdata++'s MPDATA(-FCT) advection step
 * (solvers::mpdata<ct_params_t> run under concurr::serial), used as the
 * parity oracle for the HIP library and as the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * The option bits, halo widths, advector view extents and error texts follow
 * the reference (libmpdata++/opts.hpp:34-49, concurr/detail/sharedmem.hpp:
 * 411-456,522-577, solvers/detail/solver_common.hpp:246-248,
 * solvers/detail/solver_3d.hpp:375-388).
 */
#ifndef MPDX_ORACLE_H
#define MPDX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_EQNS 64

typedef struct orc orc_t;

typedef struct {
  int n_dims;                 /* 2 or 3 */
  int n[3];                   /* grid_size */
  unsigned opts;              /* libmpdata++ opts bitmask */
  int n_iters;
  int n_eqns;
  int bc[3][2];               /* [dim][left,right]: 0 = cyclic, 1 = open */
  int upwind_filter_freq;
  int hint_scale[ORC_MAX_EQNS];
  double dt;
  double max_abs_div_eps;     /* < 0 selects numeric_limits<double>::epsilon() */
  int var_dt;                 /* ct_params_t::var_dt: adaptive time step */
  double max_courant;         /* rt_params_t::max_courant (var_dt) */
} orc_params_t;

/* returns NULL and writes a message into err on failure */
orc_t *orc_create(const orc_params_t *p, char *err, int errlen);
void orc_destroy(orc_t *s);

/* dense C-order copies of the interior advectee (n0 x n1 [x n2]) */
int orc_set_advectee(orc_t *s, int e, const double *dense);
int orc_get_advectee(const orc_t *s, int e, double *dense);

/* dense C-order copy of the reference advector(d) view: component d spans
 * face indices -1..n_d-2 along d (n_d entries); in 3D the components d>0
 * additionally span x = -1..n0-1 (n0+1 entries); other dims are interior. */
int orc_advector_extents(const orc_t *s, int d, int ext[3]);
int orc_set_advector(orc_t *s, int d, const double *dense);
int orc_get_advector(const orc_t *s, int d, double *dense);

/* runs nt more time steps; nonzero return = error (message in err) */
int orc_advance(orc_t *s, long nt, char *err, int errlen);

/* var_dt: advance by a duration (the reference's advance(real_t) with
 * ct_params_t::var_dt, solver_common.hpp:270-274,292-355, out_intrp_ord 1:
 * at least one step); dt is rescaled so that the Courant number is
 * max_courant */
int orc_advance_time(orc_t *s, double duration, char *err, int errlen);
double orc_dt(const orc_t *s);

/* time-varying advector recomputed at the start of every step (the
 * reversing_deform calc_gc pattern): kind 1 = reversing deformational flow
 * with period parameter T and largest |GC| = max_courant; kind 0 = none */
int orc_set_generator(orc_t *s, int kind, double T, double max_courant);

/* Explicit forcing of solvers::mpdata_rhs (solvers/mpdata_rhs.hpp:20-25):
 * scheme 0 none, 1 euler_a, 2 euler_b, 3 trapez.  fn plays update_rhs(dt,
 * at): it is called with the forcing arrays zeroed and writes R for every
 * equation with orc_set_rhs (reading the state with orc_get_advectee). */
typedef int (*orc_rhs_fn)(orc_t *s, double dt, int at, void *user);
int orc_set_rhs_scheme(orc_t *s, int scheme, orc_rhs_fn fn, void *user);
int orc_set_rhs(orc_t *s, int e, const double *dense);

/* Output with var_dt (output/detail/output_common.hpp:32-37,47-62,72-140):
 * records at the times k * outfreq interpolated in time (intrp_ord 1 or 2)
 * for the equations in vars.  fn(s, kind, t, user) is called with kind 0 at
 * hook_ante_loop of the first advance from timestep 0 (t = that call's end
 * time; the buffers hold the state) and kind 1 after the step completing a
 * record at time t; orc_get_output reads the buffer of equation e. */
typedef int (*orc_out_fn)(orc_t *s, int kind, double t, void *user);
int orc_set_var_dt_output(orc_t *s, double outfreq, int intrp_ord, int n_vars, const int *vars, orc_out_fn fn,
                          void *user);
int orc_get_output(const orc_t *s, int e, double *dense);

long orc_timestep(const orc_t *s);
double orc_time(const orc_t *s);
int *orc_panic_ptr(orc_t *s);

/* raw access to one full halo-extended array for white-box tests:
 * what: 0 psi[e] (current level), 1 GC[idx] */
int orc_full_shape(const orc_t *s, int *hw, int shape[3]);
int orc_get_full(const orc_t *s, int what, int idx, double *out);

#ifdef __cplusplus
}
#endif
#endif

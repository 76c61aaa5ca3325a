/*
 * HDF5 / XDMF output in the on-disk layout of libmpdata++'s output::hdf5_xdmf
 * decorator (libmpdata++/output/hdf5.hpp, output/hdf5_xdmf.hpp,
 * output/detail/xdmf_writer.hpp), so that files written from the HIP backend
 * compare with the reference's files by h5diff and open in the same XDMF
 * readers.  Host-only C ABI over the HDF5 C library (libmpdx_h5.so); the
 * arrays it takes are host copies of the solver state.
 *
 *   outdir/const.h5               X, Y(, Z): cell-corner coordinates d * index,
 *                                 float32, n+1 points per dimension;
 *                                 T: output times, float32;
 *                                 group "advection": attributes opts (string),
 *                                 dt, di, dj(, dk) (float32), var_dt (uint8),
 *                                 max_courant (float32, var_dt only),
 *                                 n_iters (int32)
 *   outdir/timestepNNNNNNNNNN.h5  one float32 dataset per output variable,
 *                                 shape = grid, one chunk, deflate level 5
 *   outdir/timestepNNNNNNNNNN.xmf XDMF grid (3DSMesh / 2DSMesh over const.h5,
 *                                 cell-centred scalar attributes)
 *   outdir/temp.xmf               temporal collection of the records
 *
 * Arrays are dense C-order (x slowest) with element size real_bytes (4: float,
 * 8: double); HDF5 converts them to float32 on write, as the reference does.
 * Every function returns 0 on success; on failure it returns nonzero and
 * mpdx_h5_last_error() describes the failure.
 */
#ifndef MPDX_H5_H
#define MPDX_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run-time description of the solver and of the output schedule
 * (output/detail/output_common.hpp:169-177 rt_params_t: outfreq, outstart,
 * outwindow; solver_common.hpp rt_params_t: dt, di/dj/dk, n_iters). */
typedef struct mpdx_h5_setup {
  int n_dims;          /* 2 or 3 */
  int grid[3];
  int real_bytes;      /* 4 or 8: the solver's real_t */
  double dijk[3];      /* di, dj, dk */
  double dt;
  int n_iters;
  uint32_t opts;       /* option bits as in include/mpdx.h (MPDX_OPT_*) */
  int var_dt;
  double max_courant;
  double outfreq;      /* steps (time with var_dt) */
  double outstart;
  int outwindow;
  int xdmf;            /* 1: output::hdf5_xdmf (write .xmf files), 0: output::hdf5 */
} mpdx_h5_setup;

/* output::hdf5::start (hdf5.hpp:66-235): creates outdir (one level) and
 * writes const.h5 for an advance over nt steps (a duration with var_dt). */
int mpdx_h5_start(const mpdx_h5_setup *s, const char *outdir, double nt);

/* output::hdf5_xdmf::record_all (hdf5_xdmf.hpp:64-85, hdf5.hpp:256-282):
 * with s->xdmf, writes timestepNNNNNNNNNN.xmf and rewrites temp.xmf over the
 * n_prev earlier records listed in prev_timesteps plus this one; then writes
 * timestepNNNNNNNNNN.h5 with one dataset per variable (names[v], fields[v]).
 * record_time is the solver's time (timestep * dt). */
int mpdx_h5_record(const mpdx_h5_setup *s, const char *outdir, int64_t timestep, double record_time,
                   int n_vars, const char *const *names, const void *const *fields,
                   int n_prev, const int64_t *prev_timesteps);

/* Number of entries of the T axis of const.h5 for an advance over nt
 * (hdf5.hpp:199-203). */
int64_t mpdx_h5_n_out(const mpdx_h5_setup *s, double nt);

/* Reads dataset `name` of the HDF5 file `path` as float32 into out (n
 * elements; fails unless the dataset holds exactly n). */
int mpdx_h5_read(const char *path, const char *name, float *out, int64_t n);

const char *mpdx_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

"""HDF5 / XDMF output of a Solver in libmpdata++'s file layout (binding of
include/mpdx_h5.h; libmpdx_h5.so is built where HDF5 is installed).

``Hdf5Xdmf`` plays the reference's ``output::hdf5_xdmf<solver>`` decorator
(output/hdf5_xdmf.hpp, output/hdf5.hpp, output/detail/output_common.hpp):
its first ``advance(nt)`` (from timestep 0) writes const.h5 and a record at
the start, as hook_ante_loop does (output_common.hpp:47-62, run only from
timestep 0 by solver_common.hpp:277-285); every advance then records after
each step that output_common's hook_post_step would record
(output_common.hpp:145-156).  With a var_dt solver the library runs the
interpolation schedule of output_common.hpp:72-140 and calls the decorator
back for the start and for each record at k * outfreq (intrp_ord plays
ct_params_t::out_intrp_ord).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB_PATH = os.path.join(HERE, "libmpdx_h5.so")
EXPORTED = ["mpdx_h5_start", "mpdx_h5_record", "mpdx_h5_n_out", "mpdx_h5_read", "mpdx_h5_last_error"]


class Setup(ctypes.Structure):
    _fields_ = [
        ("n_dims", ctypes.c_int),
        ("grid", ctypes.c_int * 3),
        ("real_bytes", ctypes.c_int),
        ("dijk", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("n_iters", ctypes.c_int),
        ("opts", ctypes.c_uint32),
        ("var_dt", ctypes.c_int),
        ("max_courant", ctypes.c_double),
        ("outfreq", ctypes.c_double),
        ("outstart", ctypes.c_double),
        ("outwindow", ctypes.c_int),
        ("xdmf", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(H5LIB_PATH):
            raise RuntimeError(f"libmpdx_h5.so not built ({H5LIB_PATH}); HDF5 output needs the HDF5 library "
                               "at build time (python -m libmpdataxx_amd.build)")
        L = ctypes.CDLL(H5LIB_PATH)
        SP = ctypes.POINTER(Setup)
        L.mpdx_h5_start.argtypes = [SP, ctypes.c_char_p, ctypes.c_double]
        L.mpdx_h5_record.argtypes = [SP, ctypes.c_char_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_void_p),
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.mpdx_h5_n_out.argtypes = [SP, ctypes.c_double]
        L.mpdx_h5_n_out.restype = ctypes.c_int64
        L.mpdx_h5_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_float),
                                   ctypes.c_int64]
        L.mpdx_h5_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc:
        raise RuntimeError(lib().mpdx_h5_last_error().decode())


def make_setup(grid, real_bytes=8, dijk=(1.0, 1.0, 1.0), dt=1.0, n_iters=2, opts=0, var_dt=False,
               max_courant=0.5, outfreq=1, outstart=0, outwindow=1, xdmf=True) -> Setup:
    s = Setup()
    s.n_dims = len(grid)
    for d, n in enumerate(grid):
        s.grid[d] = int(n)
    s.real_bytes = int(real_bytes)
    for d, v in enumerate(dijk):
        s.dijk[d] = float(v)
    s.dt = float(dt)
    s.n_iters = int(n_iters)
    s.opts = int(opts)
    s.var_dt = int(bool(var_dt))
    s.max_courant = float(max_courant)
    s.outfreq = float(outfreq)
    s.outstart = float(outstart)
    s.outwindow = int(outwindow)
    s.xdmf = int(bool(xdmf))
    return s


def start(setup: Setup, outdir, nt):
    """const.h5 for an advance over nt steps (output::hdf5::start)."""
    _check(lib().mpdx_h5_start(ctypes.byref(setup), os.fsencode(outdir), float(nt)))


def n_out(setup: Setup, nt) -> int:
    return int(lib().mpdx_h5_n_out(ctypes.byref(setup), float(nt)))


def record(setup: Setup, outdir, timestep, time, fields: dict, prev_timesteps=()):
    """timestepNNNNNNNNNN.h5 (+ .xmf and temp.xmf) with one dataset per
    {name: array} entry; arrays are converted to the setup's real type."""
    dt = np.float64 if setup.real_bytes == 8 else np.float32
    grid = tuple(setup.grid[:setup.n_dims])
    arrs = [np.ascontiguousarray(a, dtype=dt) for a in fields.values()]
    for a in arrs:
        if a.shape != grid:
            raise ValueError(f"field shape {a.shape} != grid {grid}")
    names = (ctypes.c_char_p * max(len(arrs), 1))(*[n.encode() for n in fields.keys()])
    ptrs = (ctypes.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    prev = (ctypes.c_int64 * max(len(prev_timesteps), 1))(*prev_timesteps)
    _check(lib().mpdx_h5_record(ctypes.byref(setup), os.fsencode(outdir), int(timestep), float(time), len(arrs),
                                names, ptrs, len(prev_timesteps), prev))


def read(path, name, shape) -> np.ndarray:
    """A dataset as float32 of the given shape."""
    out = np.empty(shape, dtype=np.float32)
    _check(lib().mpdx_h5_read(os.fsencode(path), name.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              out.size))
    return out


class Hdf5Xdmf:
    """output::hdf5_xdmf<solver> (output::hdf5 with xdmf=False) over a Solver.

    outvars maps equation index -> variable name (default {0: "psi"} for a
    single equation; the reference's default name is empty, which HDF5
    refuses).  dt / di / dj / dk of 0 are written as 1, as output_common's
    constructor does."""

    def __init__(self, solver, outdir, outfreq=1, outstart=0, outwindow=1, outvars=None, dijk=(0.0, 0.0, 0.0),
                 xdmf=True, intrp_ord=1):
        p = solver.params
        self.var_dt = bool(p.var_dt)
        if self.var_dt:
            # output_common.hpp:78-80,193-194
            if not outfreq > 0:
                raise ValueError("outfreq must be positive")
            if outstart != 0:
                raise ValueError("outstart must be 0 with var_dt")
            if intrp_ord not in (1, 2):
                raise ValueError("only 1st and 2nd order temporal interpolation for output implemented")
        elif outfreq <= 0 or int(outstart) % int(outfreq) != 0:
            raise ValueError("outstart must be a multiple of outfreq")
        if xdmf and solver.n_dims < 2:
            raise ValueError("only 2D and 3D output supported")
        self.solver = solver
        self.outdir = os.fspath(outdir)
        self.outvars = dict(outvars) if outvars else {0: "psi"}
        self.outstart, self.outwindow = int(outstart), int(outwindow)
        self.outfreq = float(outfreq) if self.var_dt else int(outfreq)
        self.intrp_ord = int(intrp_ord)
        dt = p.dt if p.dt != 0 else 1.0
        dijk = [v if v != 0 else 1.0 for v in list(dijk) + [0.0] * (3 - len(dijk))]
        self.setup = make_setup(solver.grid, p.real_bytes, dijk, dt, p.n_iters, p.opts, self.var_dt, p.max_courant,
                                outfreq, outstart, outwindow, xdmf)
        self._real = np.float64 if p.real_bytes == 8 else np.float32
        self.records = []
        if self.var_dt:
            solver.set_var_dt_output(self._on_output, self.outfreq, self.intrp_ord,
                                     outvars=sorted(self.outvars))

    def _on_output(self, kind, t):
        # the start (hook_ante_loop: const.h5 with the rescaled dt, a record
        # at the current time) and each interpolated record
        from . import OUT_START
        if kind == OUT_START:
            self.setup.dt = self.solver.dt
            start(self.setup, self.outdir, float(self._real(t)))
            t = self.solver.time
        ts = self.solver.timestep
        fields = {name: self.solver.output(e) for e, name in sorted(self.outvars.items())}
        record(self.setup, self.outdir, ts, float(self._real(t)), fields, self.records if self.setup.xdmf else ())
        if self.setup.xdmf:
            self.records.append(ts)

    def _time(self, ts):
        # timestep * dt in the solver's real_t (solver_common.hpp:348)
        return float(self._real(ts) * self._real(self.setup.dt))

    def record_all(self):
        ts = self.solver.timestep
        fields = {name: self.solver.advectee(e) for e, name in sorted(self.outvars.items())}
        record(self.setup, self.outdir, ts, self._time(ts), fields, self.records if self.setup.xdmf else ())
        if self.setup.xdmf:
            self.records.append(ts)

    def _records_after(self, ts):
        """Records hook_post_step makes after step ts (one per matching t)."""
        return sum(1 for t in range(self.outwindow) if (ts - t) % self.outfreq == 0 and ts - t >= self.outstart)

    def advance(self, nt):
        if self.var_dt:
            # a duration; out_intrp_ord - 1 further steps (solver_common.hpp:293-294,352)
            self.solver.advance_time(nt)
            if self.intrp_ord == 2:
                self.solver.advance(1)
            return
        if self.solver.timestep == 0:
            start(self.setup, self.outdir, nt)
            self.record_all()
        left = int(nt)
        while left > 0:
            # the steps up to the next recording one go to the library in one call
            ts = self.solver.timestep
            k = 1
            while k < left and self._records_after(ts + k) == 0:
                k += 1
            self.solver.advance(k)
            left -= k
            for _ in range(self._records_after(self.solver.timestep)):
                self.record_all()

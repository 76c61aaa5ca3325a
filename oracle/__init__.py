"""FP64 CPU oracle for the MPDATA advection step -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper around ``libmpdx_oracle.so`` (the C restatement in
``mpdx_oracle.c``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import this package, and only as the
checker / the timed CPU baseline: the product library never calls it.

Parity status of each restated function is recorded in the C file's
citations; the golden-vector and invariant tests under ``tests/`` pin them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_EQNS = 64

# libmpdata++/opts.hpp:34-49
FCT, ABS, TOT, PFC, NPA, IGA, NUG, DFL, KHN = (1 << b for b in range(9))
CYCLIC, OPEN = 0, 1
RHS_NONE, RHS_EULER_A, RHS_EULER_B, RHS_TRAPEZ = 0, 1, 2, 3
RHS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p)
OUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p)


class _Params(ctypes.Structure):
    _fields_ = [
        ("n_dims", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("opts", ctypes.c_uint),
        ("n_iters", ctypes.c_int),
        ("n_eqns", ctypes.c_int),
        ("bc", (ctypes.c_int * 2) * 3),
        ("upwind_filter_freq", ctypes.c_int),
        ("hint_scale", ctypes.c_int * MAX_EQNS),
        ("dt", ctypes.c_double),
        ("max_abs_div_eps", ctypes.c_double),
        ("var_dt", ctypes.c_int),
        ("max_courant", ctypes.c_double),
    ]


_libs: dict = {}


def build(force: bool = False) -> None:
    """Compile the oracle libraries with the committed Makefile."""
    if force or not os.path.exists(os.path.join(_HERE, "libmpdx_oracle.so")):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)


# x86-64-v4 = v3 + AVX-512 F/BW/CD/DQ/VL
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def host_isa() -> str:
    """'x86-64-v4' when this host's CPU reports the AVX-512 subsets of that
    level, else 'x86-64-v3'."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return "x86-64-v4" if _V4_FLAGS <= set(line.split(":", 1)[1].split()) else "x86-64-v3"
    except OSError:
        pass
    return "x86-64-v3"


def fast_lib_name() -> str:
    """The release-flag build that best matches this host's CPU."""
    v4 = os.path.join(_HERE, "libmpdx_oracle_fast_v4.so")
    if host_isa() == "x86-64-v4" and os.path.exists(v4):
        return "libmpdx_oracle_fast_v4.so"
    return "libmpdx_oracle_fast.so"


def _lib(fast: bool = False):
    name = fast_lib_name() if fast else "libmpdx_oracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        lib.orc_create.restype = P
        lib.orc_create.argtypes = [ctypes.POINTER(_Params), ctypes.c_char_p, ctypes.c_int]
        lib.orc_destroy.argtypes = [P]
        lib.orc_set_advectee.argtypes = [P, ctypes.c_int, dp]
        lib.orc_get_advectee.argtypes = [P, ctypes.c_int, dp]
        lib.orc_set_advector.argtypes = [P, ctypes.c_int, dp]
        lib.orc_get_advector.argtypes = [P, ctypes.c_int, dp]
        lib.orc_advector_extents.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib.orc_advance.argtypes = [P, ctypes.c_long, ctypes.c_char_p, ctypes.c_int]
        lib.orc_timestep.restype = ctypes.c_long
        lib.orc_timestep.argtypes = [P]
        lib.orc_time.restype = ctypes.c_double
        lib.orc_time.argtypes = [P]
        lib.orc_panic_ptr.restype = ctypes.POINTER(ctypes.c_int)
        lib.orc_panic_ptr.argtypes = [P]
        lib.orc_full_shape.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.orc_get_full.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        lib.orc_set_generator.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.orc_set_rhs_scheme.argtypes = [P, ctypes.c_int, RHS_FN, ctypes.c_void_p]
        lib.orc_set_rhs.argtypes = [P, ctypes.c_int, dp]
        lib.orc_advance_time.argtypes = [P, ctypes.c_double, ctypes.c_char_p, ctypes.c_int]
        lib.orc_dt.argtypes = [P]
        lib.orc_dt.restype = ctypes.c_double
        lib.orc_set_var_dt_output.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int), OUT_FN, ctypes.c_void_p]
        lib.orc_get_output.argtypes = [P, ctypes.c_int, dp]
        _libs[name] = lib
    return _libs[name]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """Mirror of ``concurr::serial<solvers::mpdata<ct>, bcond...>`` on the CPU.

    ``grid``: grid_size; ``bcs``: per-dimension (left, right) kinds;
    ``opts``: libmpdata++ option bitmask; advectee / advector arrays follow
    the reference's views (advector(d) spans faces -1..n_d-2, plus x=-1 for
    the 3D components d>0).
    """

    def __init__(self, grid, opts=FCT, n_iters=2, n_eqns=1, bcs=None,
                 dt=0.0, upwind_filter_freq=0, hint_scale=None,
                 max_abs_div_eps=-1.0, fast=False, var_dt=False, max_courant=0.5):
        self._lib = _lib(fast)
        nd = len(grid)
        if bcs is None:
            bcs = [(OPEN, OPEN)] * nd
        p = _Params()
        p.n_dims = nd
        for d in range(nd):
            p.n[d] = int(grid[d])
            p.bc[d][0], p.bc[d][1] = bcs[d]
        p.opts = int(opts)
        p.n_iters = int(n_iters)
        p.n_eqns = int(n_eqns)
        p.upwind_filter_freq = int(upwind_filter_freq)
        for e, v in enumerate(hint_scale or []):
            p.hint_scale[e] = int(v)
        p.dt = float(dt)
        p.max_abs_div_eps = float(max_abs_div_eps)
        p.var_dt = int(bool(var_dt))
        p.max_courant = float(max_courant)
        err = ctypes.create_string_buffer(256)
        h = self._lib.orc_create(ctypes.byref(p), err, 256)
        if not h:
            raise RuntimeError(err.value.decode())
        self._h = h
        self.grid = tuple(int(g) for g in grid)
        self.n_dims = nd
        self.n_eqns = n_eqns

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.orc_destroy(h)
            self._h = None

    def advector_shape(self, d):
        ext = (ctypes.c_int * 3)()
        self._lib.orc_advector_extents(self._h, d, ext)
        return tuple(ext[:self.n_dims])

    def set_advectee(self, arr, e=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.grid, (a.shape, self.grid)
        self._lib.orc_set_advectee(self._h, e, _dp(a))

    def advectee(self, e=0):
        out = np.empty(self.grid, dtype=np.float64)
        self._lib.orc_get_advectee(self._h, e, _dp(out))
        return out

    def set_advector(self, arr, d):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.advector_shape(d), (a.shape, self.advector_shape(d))
        self._lib.orc_set_advector(self._h, d, _dp(a))

    def advector(self, d):
        out = np.empty(self.advector_shape(d), dtype=np.float64)
        self._lib.orc_get_advector(self._h, d, _dp(out))
        return out

    def set_generator(self, kind, T=1.0, max_courant=0.5):
        """kind 1: reversing deformational flow recomputed every step (0: off)."""
        if self._lib.orc_set_generator(self._h, int(kind), float(T), float(max_courant)):
            raise ValueError("bad advector generator parameters")

    def set_rhs_scheme(self, scheme, fn=None):
        """scheme: RHS_EULER_A / RHS_EULER_B / RHS_TRAPEZ (0 = off); fn(dt, at)
        plays update_rhs: it sets the forcing of every equation with set_rhs
        (the arrays are zeroed before each call)."""
        def tramp(h, dt, at, user):
            try:
                if fn is not None:
                    fn(dt, at)
                return 0
            except Exception:          # noqa: BLE001 - reported as an error code
                return -1
        self._rhs_cb = RHS_FN(tramp)
        if self._lib.orc_set_rhs_scheme(self._h, int(scheme), self._rhs_cb, None):
            raise ValueError("bad rhs scheme")

    def set_rhs(self, e, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == tuple(self.grid), (a.shape, self.grid)
        if self._lib.orc_set_rhs(self._h, e, _dp(a)):
            raise ValueError("rhs not enabled or bad equation index")

    def set_var_dt_output(self, fn, outfreq, intrp_ord=1, outvars=(0,)):
        """var_dt output (output_common.hpp:47-62,72-140): fn(kind, t) with
        kind 0 at the start (t = end time of the first advance) and 1 for a
        record at time t; inside it output(e) is the field to write."""
        def tramp(h, kind, t, user):
            try:
                fn(kind, t)
                return 0
            except Exception:          # noqa: BLE001 - reported as an error code
                return -1
        self._out_cb = OUT_FN(tramp)
        vs = (ctypes.c_int * len(outvars))(*[int(e) for e in outvars])
        if self._lib.orc_set_var_dt_output(self._h, float(outfreq), int(intrp_ord), len(outvars), vs, self._out_cb,
                                           None):
            raise ValueError("bad var_dt output schedule")

    def output(self, e=0):
        out = np.empty(tuple(self.grid), dtype=np.float64)
        if self._lib.orc_get_output(self._h, e, _dp(out)):
            raise ValueError("not an output variable")
        return out

    def advance_time(self, duration):
        """var_dt: advance by a duration with the adaptive time step (at
        least one step)."""
        err = ctypes.create_string_buffer(256)
        if self._lib.orc_advance_time(self._h, float(duration), err, 256):
            raise RuntimeError(err.value.decode())

    @property
    def dt(self):
        return self._lib.orc_dt(self._h)

    def advance(self, nt):
        err = ctypes.create_string_buffer(256)
        if self._lib.orc_advance(self._h, int(nt), err, 256):
            raise RuntimeError(err.value.decode())

    @property
    def timestep(self):
        return self._lib.orc_timestep(self._h)

    @property
    def time(self):
        return self._lib.orc_time(self._h)

    def set_panic(self, v=True):
        self._lib.orc_panic_ptr(self._h)[0] = int(bool(v))

    def full(self, what, idx):
        hw = ctypes.c_int()
        sh = (ctypes.c_int * 3)()
        self._lib.orc_full_shape(self._h, ctypes.byref(hw), sh)
        shape = tuple(sh[:self.n_dims])
        out = np.empty(shape, dtype=np.float64)
        self._lib.orc_get_full(self._h, what, idx, _dp(out))
        return hw.value, out

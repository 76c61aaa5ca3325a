"""MI355X-native MPDATA advection backend (Python binding of the C ABI).

The product is the HIP library ``libmpdx.so`` (C ABI in ``include/mpdx.h``)
and the header-only C++ API in ``include/libmpdata++/``.  This module is a
thin ctypes binding used by the tests and the benchmark; it has no compute
of its own and raises if the HIP library is missing.

Names follow libmpdata++: ``Solver`` mirrors
``concurr::<backend><solvers::mpdata<ct_params_t>, bcond...>`` with
``advectee(e)``, ``advector(d)``, ``advance(nt)``, ``min/max(e)``, ``time``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpdx.so")
MAX_EQNS = 64

# option bits (libmpdata++/opts.hpp:34-49)
FCT, ABS, TOT, PFC, NPA, IGA, NUG, DFL, KHN = (1 << b for b in range(9))
CYCLIC, OPEN = 0, 1
GUARD_F64, GUARD_NATIVE, GUARD_F64_EXACT = 0, 1, 2
PATH_AUTO, PATH_PERPASS, PATH_FUSED = 0, 1, 2
RED_MIN, RED_MAX, RED_SUM = 0, 1, 2

EXPORTED = [
    "mpdx_params_default", "mpdx_create", "mpdx_destroy", "mpdx_last_error",
    "mpdx_set_stream", "mpdx_advector_extents", "mpdx_set_advectee",
    "mpdx_get_advectee", "mpdx_set_advector", "mpdx_get_advector",
    "mpdx_advance", "mpdx_reduce", "mpdx_reduce_noalloc", "mpdx_time", "mpdx_timestep",
    "mpdx_panic_ptr", "mpdx_watch_panic", "mpdx_active_path", "mpdx_z_tiles", "mpdx_profile", "mpdx_profile_read",
    "mpdx_set_calc_gc", "mpdx_advector_device", "mpdx_set_advector_generator",
    "mpdx_set_rhs_scheme", "mpdx_set_rhs", "mpdx_rhs_device", "mpdx_advectee_device",
    "mpdx_set_rhs_hooks", "mpdx_update_rhs", "mpdx_apply_rhs", "mpdx_set_var_dt_output", "mpdx_get_output",
    "mpdx_advance_time", "mpdx_dt", "mpdx_set_x_chunk", "mpdx_set_tile_y", "mpdx_tune",
    "mpdx_ensemble_create", "mpdx_ensemble_destroy", "mpdx_ensemble_advance", "mpdx_ensemble_tune",
    "mpdx_ensemble_x_chunk", "mpdx_ensemble_tile_y", "mpdx_ensemble_stats", "mpdx_ensemble_last_error",
]
GC_NONE, GC_DEFORM = 0, 1
CALC_GC_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p)
RHS_NONE, RHS_EULER_A, RHS_EULER_B, RHS_TRAPEZ, RHS_MIXED = 0, 1, 2, 3, 4
HOOK_ANTE_LOOP, HOOK_ANTE_STEP, HOOK_POST_STEP = 0, 1, 2
UPDATE_RHS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p)
RHS_HOOK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)
OUT_START, OUT_RECORD = 0, 1
OUTPUT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p)


class Params(ctypes.Structure):
    _fields_ = [
        ("n_dims", ctypes.c_int),
        ("grid", ctypes.c_int * 3),
        ("real_bytes", ctypes.c_int),
        ("opts", ctypes.c_uint32),
        ("n_iters", ctypes.c_int),
        ("n_eqns", ctypes.c_int),
        ("bc", (ctypes.c_int * 2) * 3),
        ("upwind_filter_freq", ctypes.c_int),
        ("hint_scale", ctypes.c_int * MAX_EQNS),
        ("dt", ctypes.c_double),
        ("max_abs_div_eps", ctypes.c_double),
        ("guard", ctypes.c_int),
        ("path", ctypes.c_int),
        ("var_dt", ctypes.c_int),
        ("max_courant", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libmpdx.so (build it first with ``python -m libmpdataxx_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmpdx.so not built ({LIB_PATH}); run python -m libmpdataxx_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.mpdx_params_default.argtypes = [ctypes.POINTER(Params)]
        L.mpdx_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(P)]
        L.mpdx_destroy.argtypes = [P]
        L.mpdx_last_error.restype = ctypes.c_char_p
        L.mpdx_last_error.argtypes = [P]
        L.mpdx_set_stream.argtypes = [P, P]
        L.mpdx_advector_extents.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        for name in ("mpdx_set_advectee", "mpdx_get_advectee", "mpdx_set_advector", "mpdx_get_advector",
                     "mpdx_set_rhs", "mpdx_get_output"):
            getattr(L, name).argtypes = [P, ctypes.c_int, P, i64p, ctypes.c_int]
        L.mpdx_advance.argtypes = [P, ctypes.c_int64]
        L.mpdx_advance_time.argtypes = [P, ctypes.c_double]
        L.mpdx_dt.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.mpdx_reduce.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.mpdx_reduce_noalloc.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
        L.mpdx_time.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.mpdx_timestep.argtypes = [P, i64p]
        L.mpdx_panic_ptr.restype = ctypes.POINTER(ctypes.c_int)
        L.mpdx_panic_ptr.argtypes = [P]
        L.mpdx_watch_panic.argtypes = [P, P]
        L.mpdx_active_path.argtypes = [P]
        L.mpdx_z_tiles.argtypes = [P]
        L.mpdx_set_calc_gc.argtypes = [P, CALC_GC_FN, P]
        L.mpdx_advector_device.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), i64p]
        L.mpdx_rhs_device.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), i64p]
        L.mpdx_advectee_device.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), i64p]
        L.mpdx_set_rhs_scheme.argtypes = [P, ctypes.c_int, UPDATE_RHS_FN, P]
        L.mpdx_set_rhs_hooks.argtypes = [P, RHS_HOOK_FN, P]
        L.mpdx_set_var_dt_output.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int), OUTPUT_FN, P]
        L.mpdx_update_rhs.argtypes = [P, ctypes.c_double, ctypes.c_int]
        L.mpdx_apply_rhs.argtypes = [P, ctypes.c_double]
        L.mpdx_set_advector_generator.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.mpdx_set_x_chunk.argtypes = [P, ctypes.c_int]
        L.mpdx_set_tile_y.argtypes = [P, ctypes.c_int]
        L.mpdx_tune.argtypes = [P, ctypes.c_int, i64p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.mpdx_ensemble_tile_y.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.mpdx_ensemble_stats.argtypes = [P, ctypes.POINTER(ctypes.c_int), i64p]
        L.mpdx_ensemble_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(P)]
        L.mpdx_ensemble_destroy.argtypes = [P]
        L.mpdx_ensemble_advance.argtypes = [P, ctypes.c_int64]
        L.mpdx_ensemble_tune.argtypes = [P, ctypes.c_int, i64p]
        L.mpdx_ensemble_x_chunk.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.mpdx_ensemble_last_error.restype = ctypes.c_char_p
        L.mpdx_ensemble_last_error.argtypes = [P]
        L.mpdx_profile.argtypes = [P, ctypes.c_int]
        L.mpdx_profile_read.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), i64p]
        _lib = L
    return _lib


class MpdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _strides(a: np.ndarray):
    st = [s // a.itemsize for s in a.strides] + [0] * (3 - a.ndim)
    return (ctypes.c_int64 * 3)(*st)


class Solver:
    """One MPDATA solver on the GPU (n_eqns advectees sharing one advector)."""

    def __init__(self, grid, opts=IGA | FCT, n_iters=2, n_eqns=1, bcs=None,
                 dtype=np.float64, dt=0.0, upwind_filter_freq=0, hint_scale=None,
                 max_abs_div_eps=-1.0, guard=GUARD_F64, path=PATH_AUTO, var_dt=False, max_courant=0.5):
        L = lib()
        p = Params()
        L.mpdx_params_default(ctypes.byref(p))
        nd = len(grid)
        p.n_dims = nd
        for d in range(nd):
            p.grid[d] = int(grid[d])
        if bcs is None:
            bcs = [(OPEN, OPEN)] * nd
        for d in range(nd):
            p.bc[d][0], p.bc[d][1] = bcs[d]
        self.dtype = np.dtype(dtype)
        p.real_bytes = self.dtype.itemsize
        p.opts = int(opts)
        p.n_iters = int(n_iters)
        p.n_eqns = int(n_eqns)
        p.upwind_filter_freq = int(upwind_filter_freq)
        for e, v in enumerate(hint_scale or []):
            p.hint_scale[e] = int(v)
        p.dt = float(dt)
        p.max_abs_div_eps = float(max_abs_div_eps)
        p.guard = int(guard)
        p.var_dt = int(bool(var_dt))
        p.max_courant = float(max_courant)
        p.path = int(path)
        h = ctypes.c_void_p()
        rc = L.mpdx_create(ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise MpdxError(rc, L.mpdx_last_error(None).decode())
        self._L = L
        self._h = h
        self.params = p
        self.grid = tuple(int(g) for g in grid)
        self.n_dims = nd
        self.n_eqns = n_eqns

    def close(self):
        if getattr(self, "_h", None):
            self._L.mpdx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise MpdxError(rc, self._L.mpdx_last_error(self._h).decode())

    @property
    def active_path(self):
        return self._L.mpdx_active_path(self._h)

    @property
    def z_tiles(self):
        """How the last advance() treated the z bounds in the interior
        kernel: 0 boundary shell, 1 open bounds in the tiles, 2 cyclic bounds
        through the periodic images (fields repeating plane 0 at n-1), 3 the
        same in x and y too (no boundary shell)."""
        return self._L.mpdx_z_tiles(self._h)

    def advector_shape(self, d):
        ext = (ctypes.c_int * 3)()
        self._check(self._L.mpdx_advector_extents(self._h, d, ext))
        return tuple(ext[:self.n_dims])

    # host (numpy) copies -----------------------------------------------
    def set_advectee(self, arr, e=0):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.shape == self.grid, (a.shape, self.grid)
        self._check(self._L.mpdx_set_advectee(self._h, e, a.ctypes.data, _strides(a), 0))

    def advectee(self, e=0):
        out = np.empty(self.grid, dtype=self.dtype)
        self._check(self._L.mpdx_get_advectee(self._h, e, out.ctypes.data, _strides(out), 0))
        return out

    def set_advector(self, arr, d):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.shape == self.advector_shape(d), (a.shape, self.advector_shape(d))
        self._check(self._L.mpdx_set_advector(self._h, d, a.ctypes.data, _strides(a), 0))

    def advector(self, d):
        out = np.empty(self.advector_shape(d), dtype=self.dtype)
        self._check(self._L.mpdx_get_advector(self._h, d, out.ctypes.data, _strides(out), 0))
        return out

    # device (torch) copies ----------------------------------------------
    def _device_view(self, t, shape, what):
        """Element strides of a device tensor, after checking that the kernel
        that copies it stays inside it: the shape of the view, the element
        type of the solver, memory on a GPU and non-negative strides.  Raises
        ValueError before anything is launched."""
        if not getattr(t, "is_cuda", False):
            raise ValueError(f"{what}: expected a GPU tensor")
        want = {4: "torch.float32", 8: "torch.float64"}[self.dtype.itemsize]
        if str(t.dtype) != want:
            raise ValueError(f"{what}: dtype {t.dtype}, the solver's is {want}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{what}: shape {tuple(t.shape)}, expected {tuple(shape)}")
        if any(s < 0 for s in t.stride()):
            raise ValueError(f"{what}: negative strides")
        return (ctypes.c_int64 * 3)(*(list(t.stride()) + [0] * (3 - t.dim())))

    def set_advectee_device(self, t, e=0):
        st = self._device_view(t, self.grid, "advectee")
        self._check(self._L.mpdx_set_advectee(self._h, e, t.data_ptr(), st, 1))

    def get_advectee_device(self, t, e=0):
        st = self._device_view(t, self.grid, "advectee")
        self._check(self._L.mpdx_get_advectee(self._h, e, t.data_ptr(), st, 1))

    def set_advector_device(self, t, d):
        st = self._device_view(t, self.advector_shape(d), f"advector({d})")
        self._check(self._L.mpdx_set_advector(self._h, d, t.data_ptr(), st, 1))

    # stepping and diagnostics ---------------------------------------------
    def set_stream(self, stream_ptr):
        self._check(self._L.mpdx_set_stream(self._h, stream_ptr))

    def advance(self, nt):
        self._step_check(self._L.mpdx_advance(self._h, int(nt)))

    def _step_check(self, rc):
        # an exception raised by an output callback during the call
        exc, self._out_exc = getattr(self, "_out_exc", None), None
        if rc and exc is not None:
            raise exc
        self._check(rc)

    def reduce(self, op, e=0):
        out = ctypes.c_double()
        self._check(self._L.mpdx_reduce(self._h, e, op, ctypes.byref(out)))
        return out.value

    def min(self, e=0):
        return self.reduce(RED_MIN, e)

    def max(self, e=0):
        return self.reduce(RED_MAX, e)

    def mass(self, e=0):
        """FP64 Kahan-compensated sum of the advectee over the grid."""
        return self.reduce(RED_SUM, e)

    @property
    def time(self):
        t = ctypes.c_double()
        self._L.mpdx_time(self._h, ctypes.byref(t))
        return t.value

    def advance_time(self, duration):
        """var_dt solvers: advance by a duration with the adaptive step.

        Steps while time < time at the call + duration, and at least one
        (advance(real_t), solver_common.hpp:292-355, with out_intrp_ord 1):
        advance_time(0) takes one step.  advance(n) takes n steps of the
        current dt."""
        self._step_check(self._L.mpdx_advance_time(self._h, float(duration)))

    def set_var_dt_output(self, fn, outfreq, intrp_ord=1, outvars=(0,)):
        """var_dt output (output_common.hpp:47-62,72-140): fn(kind, t) is
        called with OUT_START (first advance from timestep 0, t = the call's
        end time) and OUT_RECORD (a record at time t = k * outfreq,
        interpolated to order intrp_ord); inside it output(e) is the field to
        write.  fn = None removes the schedule."""
        if fn is None:
            self._out_cb = None
            self._check(self._L.mpdx_set_var_dt_output(self._h, 1.0, 1, 0, None, OUTPUT_FN(), None))
            return

        def tramp(h, kind, t, user):
            try:
                fn(kind, t)
                return 0
            except Exception as exc:   # noqa: BLE001 - re-raised by the advance call
                self._out_exc = exc
                return -1
        vs = (ctypes.c_int * len(outvars))(*[int(e) for e in outvars])
        cb = OUTPUT_FN(tramp)
        self._check(self._L.mpdx_set_var_dt_output(self._h, float(outfreq), int(intrp_ord), len(outvars), vs, cb,
                                                   None))
        self._out_cb = cb

    def output(self, e=0):
        """The interpolation buffer of output equation e (interior)."""
        out = np.empty(self.grid, dtype=self.dtype)
        self._check(self._L.mpdx_get_output(self._h, e, out.ctypes.data, _strides(out), 0))
        return out

    @property
    def dt(self):
        t = ctypes.c_double()
        self._L.mpdx_dt(self._h, ctypes.byref(t))
        return t.value

    @property
    def timestep(self):
        t = ctypes.c_int64()
        self._L.mpdx_timestep(self._h, ctypes.byref(t))
        return t.value

    def profile(self, enable=True):
        """Start (clearing) or stop per-kernel-class HIP-event timing."""
        self._check(self._L.mpdx_profile(self._h, int(bool(enable))))

    def profile_read(self):
        """{kernel class: (total device ms, launches)} recorded so far."""
        out = {}
        idx = 0
        buf = ctypes.create_string_buffer(128)
        while True:
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            if self._L.mpdx_profile_read(self._h, idx, buf, 128, ctypes.byref(ms), ctypes.byref(n)):
                break
            out[buf.value.decode()] = (ms.value, n.value)
            idx += 1
        return out

    def set_calc_gc(self, fn):
        """fn(t_half) -> bool, called once per step like the reference's
        calc_gc(); it may call set_advector / set_advector_device."""
        if fn is None:
            self._calc_gc = None
            self._check(self._L.mpdx_set_calc_gc(self._h, CALC_GC_FN(), None))
            return

        def tramp(h, t_half, user):
            try:
                return 1 if fn(t_half) else 0
            except Exception:          # noqa: BLE001 - reported as an error code
                return -1
        self._calc_gc = CALC_GC_FN(tramp)
        self._check(self._L.mpdx_set_calc_gc(self._h, self._calc_gc, None))

    def set_rhs_scheme(self, scheme, fn=None):
        """Explicit forcing (solvers::mpdata_rhs): scheme RHS_EULER_A /
        RHS_EULER_B / RHS_TRAPEZ (RHS_NONE = off); fn(dt, at) plays
        update_rhs and sets each equation's forcing with set_rhs (the forcing
        is zeroed before each call)."""
        def tramp(h, dt, at, user):
            try:
                if fn is not None:
                    fn(dt, at)
                return 0
            except Exception:          # noqa: BLE001 - reported as an error code
                return -1
        self._rhs_cb = UPDATE_RHS_FN(tramp)
        self._check(self._L.mpdx_set_rhs_scheme(self._h, int(scheme), self._rhs_cb, None))

    def set_rhs_hooks(self, hook):
        """The mixed scheme's hooks: hook(which) with which = HOOK_ANTE_LOOP,
        HOOK_ANTE_STEP or HOOK_POST_STEP; it calls update_rhs / apply_rhs."""
        def tramp(h, which, user):
            try:
                hook(which)
                return 0
            except Exception:          # noqa: BLE001 - reported as an error code
                return -1
        self._rhs_hook = RHS_HOOK_FN(tramp)
        self._check(self._L.mpdx_set_rhs_hooks(self._h, self._rhs_hook, None))

    def update_rhs(self, dt, at):
        self._check(self._L.mpdx_update_rhs(self._h, float(dt), int(at)))

    def apply_rhs(self, dt):
        self._check(self._L.mpdx_apply_rhs(self._h, float(dt)))

    def set_rhs(self, e, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.shape == tuple(self.grid), (a.shape, self.grid)
        self._check(self._L.mpdx_set_rhs(self._h, e, a.ctypes.data, _strides(a), 0))

    def set_advector_generator(self, kind, *params):
        arr = (ctypes.c_double * max(len(params), 1))(*params)
        self._check(self._L.mpdx_set_advector_generator(self._h, kind, arr, len(params)))

    def set_panic(self, v=True):
        self._L.mpdx_panic_ptr(self._h)[0] = int(bool(v))

    def set_x_chunk(self, planes):
        """x-chunk length of the 3D interior kernel (0 = automatic)."""
        self._check(self._L.mpdx_set_x_chunk(self._h, int(planes)))

    def set_tile_y(self, rows):
        """Interior tile height (0 = default, 16 or 32; FP32 default guards)."""
        self._check(self._L.mpdx_set_tile_y(self._h, int(rows)))

    def tune(self, steps_per_trial=3):
        """Autotune the interior tile height and x-chunk length for this
        grid (mpdx_tune); the trials are real steps.  Returns (steps taken,
        tile height kept, x chunk kept)."""
        n, ty, xc = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.mpdx_tune(self._h, int(steps_per_trial), ctypes.byref(n), ctypes.byref(ty),
                                      ctypes.byref(xc)))
        return n.value, ty.value, xc.value


class Ensemble:
    """Independent solvers advanced together: each step is one graph of
    batched launches over all members (mpdx_ensemble_*); the members stay
    usable on their own."""

    def __init__(self, members):
        L = lib()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        rc = L.mpdx_ensemble_create(arr, len(self.members), ctypes.byref(h))
        if rc:
            raise MpdxError(rc, L.mpdx_ensemble_last_error(None).decode())
        self._L = L
        self._h = h

    def _check(self, rc):
        if rc:
            raise MpdxError(rc, self._L.mpdx_ensemble_last_error(self._h).decode())

    def advance(self, nt):
        self._check(self._L.mpdx_ensemble_advance(self._h, int(nt)))

    def tune(self, steps_per_trial=2):
        """Autotune the x-chunk length per class of 3D members; returns the
        number of (real) steps the trials took."""
        n = ctypes.c_int64()
        self._check(self._L.mpdx_ensemble_tune(self._h, int(steps_per_trial), ctypes.byref(n)))
        return n.value

    def x_chunk(self, i):
        v = ctypes.c_int()
        self._check(self._L.mpdx_ensemble_x_chunk(self._h, int(i), ctypes.byref(v)))
        return v.value

    def tile_y(self, i):
        v = ctypes.c_int()
        self._check(self._L.mpdx_ensemble_tile_y(self._h, int(i), ctypes.byref(v)))
        return v.value

    def stats(self):
        """(batched launches, tasks) of one ensemble step."""
        n, t = ctypes.c_int(), ctypes.c_int64()
        self._check(self._L.mpdx_ensemble_stats(self._h, ctypes.byref(n), ctypes.byref(t)))
        return n.value, t.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.mpdx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/mpdx.h declares, the parameter defaults match the reference's
ct_params_default_t / rt_params_t, and the package refuses to run without the
HIP library (no CPU fallback)."""
import ctypes
import os
import re

import pytest

import libmpdataxx_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mpdx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpdx_[a-z_]+)\s*\(", text)))


def test_header_declares_expected_api():
    assert declared_symbols() == sorted(M.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(M.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_params_default_matches_reference():
    p = M.Params()
    M.lib().mpdx_params_default(ctypes.byref(p))
    assert p.opts == M.IGA | M.FCT          # opts.hpp:83
    assert p.n_iters == 2                   # mpdata_common.hpp:76
    assert p.n_eqns == 1 and p.max_abs_div_eps < 0
    assert p.guard == M.GUARD_F64 and p.path == M.PATH_AUTO


@pytest.mark.parametrize("kw,msg", [
    (dict(grid=(0, 8)), "bogus grid size"),
    (dict(grid=(8, 8), opts=M.ABS | M.IGA), "mutually exclusive"),
    (dict(grid=(8, 8), opts=M.TOT), "not supported"),
    (dict(grid=(8, 8), opts=M.IGA, n_iters=3), "one corrective step"),
])
def test_create_rejects_like_reference(kw, msg):
    with pytest.raises(M.MpdxError, match=msg):
        M.Solver(**kw)


def test_null_handles_are_errors():
    """Every entry point that takes a solver or ensemble handle returns
    MPDX_EINVAL for a null one (the check comes before any HIP call), with
    the message from mpdx_last_error(NULL)."""
    L = M.lib()
    d, i64 = ctypes.c_double(), ctypes.c_int64()
    ext = (ctypes.c_int * 3)()
    st = (ctypes.c_int64 * 3)()
    buf = (ctypes.c_double * 8)()
    calls = {
        "mpdx_set_stream": (None, None),
        "mpdx_advector_extents": (None, 0, ext),
        "mpdx_set_advectee": (None, 0, buf, st, 0),
        "mpdx_get_advectee": (None, 0, buf, st, 0),
        "mpdx_set_advector": (None, 0, buf, st, 0),
        "mpdx_get_advector": (None, 0, buf, st, 0),
        "mpdx_advance": (None, 1),
        "mpdx_advance_time": (None, 1.0),
        "mpdx_dt": (None, ctypes.byref(d)),
        "mpdx_reduce": (None, 0, 0, ctypes.byref(d)),
        "mpdx_time": (None, ctypes.byref(d)),
        "mpdx_timestep": (None, ctypes.byref(i64)),
        "mpdx_set_x_chunk": (None, 0),
        "mpdx_set_tile_y": (None, 0),
        "mpdx_tune": (None, 1, None, None, None),
        "mpdx_profile": (None, 1),
        "mpdx_profile_read": (None, 0, None, 0, None, None),
        "mpdx_set_calc_gc": (None, M.CALC_GC_FN(), None),
        "mpdx_set_rhs_scheme": (None, 0, M.UPDATE_RHS_FN(), None),
        "mpdx_set_rhs": (None, 0, buf, st, 0),
        "mpdx_set_advector_generator": (None, 1, None, 0),
        "mpdx_advector_device": (None, 0, ctypes.byref(ctypes.c_void_p()), st),
        "mpdx_advectee_device": (None, 0, ctypes.byref(ctypes.c_void_p()), st),
        "mpdx_rhs_device": (None, 0, ctypes.byref(ctypes.c_void_p()), st),
        "mpdx_ensemble_advance": (None, 1),
        "mpdx_ensemble_tune": (None, 1, None),
        "mpdx_ensemble_x_chunk": (None, 0, ctypes.byref(ctypes.c_int())),
        "mpdx_set_var_dt_output": (None, 1.0, 1, 1, None, M.OUTPUT_FN(), None),
        "mpdx_get_output": (None, 0, buf, st, 0),
    }
    for name, args in calls.items():
        fn = getattr(L, name)
        assert fn(*args) == 1, name                       # MPDX_EINVAL
        assert b"null" in L.mpdx_last_error(None), name
    assert L.mpdx_active_path(None) == -1
    assert not L.mpdx_panic_ptr(None)                     # a null pointer, not a crash
    assert b"null" in L.mpdx_last_error(None)
    p = M.Params()
    L.mpdx_params_default(ctypes.byref(p))
    assert L.mpdx_create(ctypes.byref(p), None) == 1


def test_null_buffers_are_errors():
    """Null buffers, strides and output pointers are rejected before the
    handle is used: a stand-in handle (host memory that no call may
    dereference) with a null pointer argument gives MPDX_EINVAL."""
    L = M.lib()
    fake = ctypes.create_string_buffer(256)
    h = ctypes.cast(fake, ctypes.c_void_p)
    st = (ctypes.c_int64 * 3)(1, 1, 1)
    buf = (ctypes.c_double * 8)()
    pp = ctypes.byref(ctypes.c_void_p())
    calls = {
        "mpdx_set_advectee": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_get_advectee": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_set_advector": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_get_advector": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_set_rhs": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_advector_device": [(h, 0, None, st), (h, 0, pp, None)],
        "mpdx_advectee_device": [(h, 0, None, st), (h, 0, pp, None)],
        "mpdx_rhs_device": [(h, 0, None, st), (h, 0, pp, None)],
        "mpdx_dt": [(h, None)],
        "mpdx_time": [(h, None)],
        "mpdx_timestep": [(h, None)],
        "mpdx_reduce": [(h, 0, 0, None)],
        "mpdx_advector_extents": [(h, 0, None)],
        "mpdx_get_output": [(h, 0, None, st, 0), (h, 0, buf, None, 0)],
        "mpdx_set_var_dt_output": [(h, 1.0, 1, 1, None, M.OUTPUT_FN(lambda *a: 0), None)],
    }
    for name, variants in calls.items():
        fn = getattr(L, name)
        for args in variants:
            assert fn(*args) == 1, (name, args)
            assert b"null" in L.mpdx_last_error(None), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(M, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(M, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        M.lib()


def test_box_index_fast_division(tmp_path):
    """The box kernels split flat plane indices with a multiply-high division
    by the box's run-time extent; its host half is compiled with g++ here and
    checked against exact integer division over divisors 1..4096, powers of
    two and their neighbours and large odd values, for dividends up to 2^31-1."""
    import subprocess
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "libmpdataxx_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "fastdiv_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


def test_box_list_launch_covers_every_cell_once(tmp_path):
    """The halo and vector-fill kernels run as box lists: up to 16 boxes of
    different shapes in one 1D grid of 256-thread blocks, each thread
    decoding its (box, a, b, c) from a block prefix table and two fast
    divisions.  The host half of that decode is compiled with g++ and every
    thread of 150 random lists (overlapping, empty and one-cell boxes, more
    than 16 boxes per list) is enumerated: each cell of each box must come
    out exactly once.  A box of more than 2^31 cells must be cut into
    contiguous pieces below 2^31 cells along its first coordinate."""
    import subprocess
    exe = str(tmp_path / "boxlist_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "libmpdataxx_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "boxlist_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


def test_interior_kernels_do_not_spill():
    """The x-marching interior kernels (single and batched launches) run one
    1024-thread block (or two 512-thread ones) per CU with at most 128 VGPRs
    per thread; a register spill to scratch memory costs them up to 2x.  The build records the compiler's per-kernel resource
    usage; every interior kernel variant must have no VGPR spills and no
    scratch."""
    path = os.path.join(ROOT, "libmpdataxx_amd", "kernel_resources.txt")
    if not os.path.exists(path):
        pytest.skip("library not built here")
    rows = [l.split() for l in open(path) if "k_fz_interior" in l or "InteriorTask" in l]
    assert rows, "no interior kernels in the resource report"
    for r in rows:
        f = dict(kv.split("=") for kv in r[1:])
        assert f["vspill"] == "0" and f["scratch"] == "0", r


class _NoCalls:
    """Stands in for the C library: any call is a test failure."""

    def __getattr__(self, name):
        raise AssertionError(f"{name} called although the arguments are invalid")


class _FakeTensor:
    def __init__(self, shape, dtype="torch.float32", cuda=True, stride=None):
        self.shape, self.dtype, self.is_cuda = tuple(shape), dtype, cuda
        st, s = [], 1
        for n in reversed(shape):
            st.insert(0, s)
            s *= n
        self._stride = tuple(stride) if stride else tuple(st)

    def stride(self):
        return self._stride

    def dim(self):
        return len(self.shape)

    def data_ptr(self):
        return 4096


@pytest.mark.parametrize("t,msg", [
    (_FakeTensor((6, 5, 4), cuda=False), "GPU tensor"),
    (_FakeTensor((6, 5, 4), dtype="torch.float64"), "dtype"),
    (_FakeTensor((6, 5, 3)), "shape"),
    (_FakeTensor((5, 5, 4)), "shape"),
    (_FakeTensor((6, 5, 4), stride=(20, 4, -1)), "negative"),
])
def test_device_views_are_checked_before_launch(t, msg):
    """set/get_advectee_device and set_advector_device pass a tensor's raw
    address and strides to a copy kernel; a tensor of the wrong shape,
    element type or memory would make that kernel read or write outside it,
    so the binding rejects it before calling the library."""
    import numpy as np
    s = M.Solver.__new__(M.Solver)
    s._L, s._h = _NoCalls(), None
    s.grid, s.dtype, s.n_dims = (6, 5, 4), np.dtype(np.float32), 3
    s.advector_shape = lambda d: (6, 5, 4)
    for call in (lambda: s.set_advectee_device(t), lambda: s.get_advectee_device(t),
                 lambda: s.set_advector_device(t, 0)):
        with pytest.raises(ValueError, match=msg):
            call()

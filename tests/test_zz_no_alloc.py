"""mpdx_reduce_noalloc against mpdx_reduce, through the C ABI: the same
kernels with the block partials in the library's fixed per-GPU array and the
result written by the GPU to a caller's device buffer.  Two calls in a row on
different solvers, the second on a smaller grid (the array keeps the first
call's partials, which the second must overwrite), for min, max and the FP64
compensated sum, FP32 and FP64, 2D and 3D: bit-identical results.  A grid
with more x planes than the array holds is refused before any launch."""
import ctypes

import numpy as np
import pytest

import libmpdataxx_amd as M
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(s, op, out):
    rc = s._L.mpdx_reduce_noalloc(s._h, 0, op, ctypes.c_void_p(out.data_ptr()))
    assert rc == 0, s._L.mpdx_last_error(s._h)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("grids", [[(40, 34, 20), (20, 18, 12)], [(300, 70), (70, 33)]])
def test_reduce_noalloc_matches_reduce(dtype, grids):
    import torch
    out = torch.full((4,), np.nan, dtype=torch.float64, device="cuda:0")
    for n, grid in enumerate(grids):
        print(f"reduce_noalloc grid={grid} dtype={np.dtype(dtype).name}", flush=True)
        s = M.Solver(grid, opts=M.FCT, n_iters=2, dtype=dtype, bcs=[(M.CYCLIC, M.CYCLIC)] * len(grid))
        s.set_advectee(W.random_field(grid, 20 + n, lo=-1.0, hi=3.0))
        for op in (M.RED_MIN, M.RED_MAX, M.RED_SUM):
            want = s.reduce(op)
            o = out[op:op + 1]
            _noalloc(s, op, o)
            torch.cuda.synchronize()
            got = float(o.item())
            assert got == want, (grid, op, got, want)
        s.close()


def test_reduce_noalloc_refuses_too_many_planes():
    import torch
    out = torch.zeros((1,), dtype=torch.float64, device="cuda:0")
    s = M.Solver((9000, 12, 12), opts=0, n_iters=1, dtype=np.float32, bcs=[(M.CYCLIC, M.CYCLIC)] * 3)
    rc = s._L.mpdx_reduce_noalloc(s._h, 0, M.RED_SUM, ctypes.c_void_p(out.data_ptr()))
    assert rc == 1
    assert b"8192 x planes" in s._L.mpdx_last_error(s._h)
    s.close()

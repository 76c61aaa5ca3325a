"""Device-pointer copies of the C ABI (mpdx_set/get_advectee, _advector with
on_device = 1): strided views, and ordering against work that the caller
enqueued on another stream (torch's) just before the call."""
import numpy as np
import pytest
import torch

import libmpdataxx_amd as M
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_strided_views_round_trip(dtype):
    grid = (12, 10, 14)
    print(f"strided views dtype={dtype}", flush=True)
    s = M.Solver(grid, opts=M.FCT, n_iters=2, dtype=np.float32 if dtype == torch.float32 else np.float64)
    base = torch.rand(grid[2], grid[1], grid[0], dtype=dtype, device=DEV)
    view = base.permute(2, 1, 0)                      # [x][y][z] view of a z-major array
    s.set_advectee_device(view)
    np.testing.assert_array_equal(s.advectee(), view.cpu().numpy())
    out = torch.full((grid[0], grid[1] * 2, grid[2]), -1.0, dtype=dtype, device=DEV)[:, ::2, :]
    s.get_advectee_device(out)
    np.testing.assert_array_equal(out.cpu().numpy(), view.cpu().numpy())
    for d in range(3):
        ext = s.advector_shape(d)
        a = torch.rand(ext, dtype=dtype, device=DEV)
        s.set_advector_device(a, d)
        np.testing.assert_array_equal(s.advector(d), a.cpu().numpy())
    s.close()


def test_copy_waits_for_producer_on_another_stream():
    """The library's stream is non-blocking: a field still being written by
    a kernel on torch's stream when set_advectee_device is called must be
    copied complete, and a torch kernel issued right after
    get_advectee_device must read the copied values."""
    grid = (256, 256, 256)
    print("producer ordering 256^3", flush=True)
    s = M.Solver(grid, opts=M.FCT, n_iters=2, dtype=np.float32)
    x = torch.rand(grid, dtype=torch.float32, device=DEV)
    torch.cuda.synchronize()
    for rep in range(3):
        y = x
        for _ in range(20):                           # a long chain of kernels on torch's stream
            y = torch.sin(y) * 0.5 + 0.5
        s.set_advectee_device(y)                      # no synchronisation before the call
        got = torch.empty(grid, dtype=torch.float32, device=DEV)
        s.get_advectee_device(got)
        d = (got - y).abs().max()                     # consumer on torch's stream
        assert float(d) == 0.0, rep
        x = y
    # the zero advector produced just before the first advance passes the
    # divergence check (it is all zeros, not a field still being written)
    for d in range(3):
        s.set_advector_device(torch.zeros(s.advector_shape(d), dtype=torch.float32, device=DEV), d)
    s.advance(1)
    s.close()


def test_state_unchanged_by_zero_advector_after_device_copies():
    grid = (24, 20, 22)
    psi = W.random_field(grid, 3)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, dtype=np.float64)
    s.set_advectee_device(torch.from_numpy(psi).to(DEV))
    for d in range(3):
        s.set_advector_device(torch.zeros(s.advector_shape(d), dtype=torch.float64, device=DEV), d)
    s.advance(2)
    np.testing.assert_array_equal(s.advectee(), psi)
    s.close()

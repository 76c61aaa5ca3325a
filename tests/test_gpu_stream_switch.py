"""mpdx_set_stream (concurr/any.hpp's solver bound to a caller's queue):
switching a solver from its own stream to a caller's stream and back keeps
every step's result, including replays of the step graphs captured before
the switch (a graph is launched on whatever stream the solver holds)."""
import numpy as np
import pytest
import torch

import libmpdataxx_amd as M
from gpu_util import setup_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stream_switch_keeps_results(dtype):
    grid = (40, 34, 30)
    bcs = [(M.OPEN, M.OPEN)] * 3
    psi, gc = setup_case(grid, bcs, 7)
    print(f"stream switch grid={grid} dtype={np.dtype(dtype).name}", flush=True)

    def make():
        s = M.Solver(grid, opts=M.FCT, n_iters=3, bcs=bcs, dtype=dtype, max_abs_div_eps=1e300)
        s.set_advectee(psi)
        for d in range(3):
            s.set_advector(gc[d], d)
        return s

    ref = make()
    ref.advance(9)
    want = ref.advectee()
    ref.close()

    s = make()
    s.advance(3)                      # graphs captured on the solver's own stream
    caller = torch.cuda.Stream()      # default priority
    s.set_stream(caller.cuda_stream)
    s.advance(3)
    s.set_stream(None)                # a new stream of its own
    s.advance(3)
    got = s.advectee()
    s.close()
    np.testing.assert_array_equal(got, want)

"""The single-solver autotuner (mpdx_tune): it times the interior kernel's
tile heights and x-chunk lengths on real steps and keeps the fastest pair.
The trial steps are ordinary steps, so a tuned solver must hold exactly what
an untuned one holds after the same number of steps (results do not depend
on the pair), and in FP64 match the oracle after them."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu


def _sphere(n, dtype, bcs=None):
    psi, gc = W.rotation_sphere_3d(n)
    kw = {} if bcs is None else dict(bcs=bcs)
    s = M.Solver((n,) * 3, opts=M.FCT, n_iters=3, dtype=dtype, **kw)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    return s, psi, gc


@pytest.mark.parametrize("n", [40, 64])
def test_tuned_fp32_equals_untuned(n):
    print(f"tune FP32 {n}^3 FCT n_iters=3 open", flush=True)
    s, _, _ = _sphere(n, np.float32)
    taken, ty, xc = s.tune(2)
    # 3 tile heights x the chunks up to the x extent rounded up to 2^k, 2 + 2
    # steps each
    span = 16
    while span < n - 4:
        span *= 2
    chunks = [c for c in (0, 16, 32, 64, 128, 256) if c <= span]
    assert taken == 3 * len(chunks) * 4
    assert ty in (0, 32, 16) and xc in chunks
    assert s.timestep == taken
    s.advance(5)
    ref, _, _ = _sphere(n, np.float32)
    ref.advance(taken + 5)
    np.testing.assert_array_equal(s.advectee(), ref.advectee())
    s.close()
    ref.close()


def test_tuned_fp64_matches_oracle():
    n = 36
    print(f"tune FP64 {n}^3 FCT n_iters=3 cyclic vs oracle", flush=True)
    bcs = [(M.CYCLIC, M.CYCLIC)] * 3
    s, psi, gc = _sphere(n, np.float64, bcs)
    taken, ty, xc = s.tune(1)
    assert taken == 3 * 3     # FP64: the default tile height, chunks 0, 16, 32
    assert ty == 0 and xc in (0, 16, 32)
    s.advance(3)
    o = O.Oracle((n,) * 3, opts=O.FCT, n_iters=3, bcs=bcs)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(taken + 3)
    want = o.advectee()
    assert np.abs(s.advectee() - want).max() <= 1e-12 * np.abs(want).max()
    s.close()


def test_tune_without_interior_kernel_takes_no_steps():
    print("tune 2D solver (no 3D interior kernel)", flush=True)
    s = M.Solver((64, 48), opts=M.FCT, n_iters=2, dtype=np.float32)
    s.set_advectee(np.ones((64, 48)))
    for d in range(2):
        s.set_advector(np.full(s.advector_shape(d), 0.1), d)
    assert s.tune(3) == (0, 0, 0)
    assert s.timestep == 0
    with pytest.raises(M.MpdxError):
        s.tune(0)
    s.close()

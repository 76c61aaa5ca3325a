"""z-boundary tiles: with open bounds on both z sides the interior kernel's
tiles also compute the cells next to the z boundaries (scalar halo columns
from the edge copy, the velocity halo from the open boundary's normal-fill
zero and zero-divergence rules), and the boundary shell has no z slabs.

In FP64 the results must match the oracle for every option set of the
interior kernel and z extents where one tile holds both boundaries (12, 27,
28) or the last tile overlaps its neighbour (29, 45, 61), with cyclic or
open x and y bounds.  Against the boundary-shell path (MPDX_NO_ZBND=1) they may
differ only by rounding: the two kernels compute a z-boundary cell with the
same formulae but the compiler contracts products into fused multiply-adds
differently in each."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu

OPTS = [(M.FCT, 2), (M.FCT, 3), (M.FCT, 4), (0, 2), (0, 3), (M.IGA, 2), (M.IGA | M.FCT, 2)]


def run(grid, opts, iters, dtype, bcs, psi, gc, steps, monkeypatch, zbnd):
    if zbnd:
        monkeypatch.delenv("MPDX_NO_ZBND", raising=False)
    else:
        monkeypatch.setenv("MPDX_NO_ZBND", "1")
    s = M.Solver(grid, opts=opts, n_iters=iters, dtype=dtype, bcs=bcs)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    s.advance(steps)
    out = s.advectee()
    s.close()
    return out


NZ = [12, 27, 28, 29, 45, 61]


def case(nz, seed):
    grid = (22, 19, nz)
    xb = M.CYCLIC if nz % 2 else M.OPEN
    yb = M.CYCLIC if nz % 3 == 0 and xb == M.CYCLIC else M.OPEN
    bcs = [(xb, xb), (yb, yb), (M.OPEN, M.OPEN)]
    psi = W.random_field(grid, seed + nz)
    gc = W.streamfunction_advector(grid, seed + 100 + nz, periodic=(xb == M.CYCLIC, yb == M.CYCLIC, False),
                                   max_courant=0.45)
    return grid, bcs, psi, gc


@pytest.mark.parametrize("nz", NZ)
@pytest.mark.parametrize("opts,iters", OPTS)
def test_zbnd_tiles_match_oracle_fp64(nz, opts, iters, monkeypatch):
    grid, bcs, psi, gc = case(nz, 100)
    print(f"zbnd vs oracle nz={nz} opts={opts} n_iters={iters} x/y bounds={bcs[0][0]}/{bcs[1][0]}", flush=True)
    got = run(grid, opts, iters, np.float64, bcs, psi, gc, 4, monkeypatch, True)
    oo = (O.FCT if opts & M.FCT else 0) | (O.IGA if opts & M.IGA else 0)
    ob = [(O.CYCLIC if b[0] == M.CYCLIC else O.OPEN,) * 2 for b in bcs]
    o = O.Oracle(grid, opts=oo, n_iters=iters, bcs=ob)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(4)
    want = o.advectee()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("nz", NZ)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("opts,iters", OPTS)
def test_zbnd_tiles_round_like_shell_path(nz, dtype, opts, iters, monkeypatch):
    grid, bcs, psi, gc = case(nz, 300)
    print(f"zbnd vs shell nz={nz} dtype={np.dtype(dtype).name} opts={opts} n_iters={iters}", flush=True)
    a = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, True)
    b = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, False)
    eps = np.finfo(dtype).eps
    assert np.abs(a.astype(np.float64) - b).max() <= 16 * eps * np.abs(b).max()


@pytest.mark.parametrize("iters,scales", [(3, [0, 3]), (4, [-2, 1])])
def test_zbnd_tiles_two_equations_hint_scale_fp64(iters, scales):
    """Two equations with their own hint_scale (the tiles' scalar loads and
    the halo columns see the scaled values) on a grid that takes the
    interior kernel, against the oracle."""
    from gpu_util import rel_err, run_gpu, run_oracle, setup_case
    grid, bcs = (20, 18, 30), [(1, 1)] * 3
    print(f"zbnd two equations n_iters={iters} hint_scale={scales}", flush=True)
    psi, gc = setup_case(grid, bcs, 41)
    psi2 = psi[::-1, :, ::-1].copy()
    want = run_oracle(grid, bcs, O.FCT, iters, [psi, psi2], gc, 4, n_eqns=2, hint_scale=scales)
    got = run_gpu(grid, bcs, O.FCT, iters, [psi, psi2], gc, 4, n_eqns=2, hint_scale=scales, path=2)
    for g, w in zip(got, want):
        assert rel_err(g, w) <= 1e-12

"""The 2D pass kernel (k_fz_pass2d): one launch per corrective pass, the first
one computing psi^1 itself from psi^n and GC (formulae_mpdata_2d.hpp:241-278,
formulae_mpdata_fct_2d.hpp:21-120, donorcell_formulae.hpp:50-79).

Open or mixed bounds: the kernel covers the cells 2..n-3 next to the
boundary-shell kernels.  Cyclic bounds on both sides of x and y with fields
repeating plane 0 at n-1 along both (z_tiles 3): it covers every cell, reading
frame cells at their periodic images, with no shell and no halo fill.  In FP64
the results must match the oracle; against the shell kernels over every cell
(MPDX_NO_PASS2D=1) they may differ only by rounding (the compiler contracts
products into fused multiply-adds differently in each kernel)."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu

OPTS = [(M.FCT, 2), (M.FCT, 3), (M.FCT, 4), (0, 2), (0, 3), (M.IGA, 2), (M.IGA | M.FCT, 2)]
SIZES = [(12, 12), (27, 45), (64, 33), (100, 130)]
BOUNDS = ["cyclic", "open", "mixed"]


def case2d(grid, seed, kind):
    xb = M.OPEN if kind == "open" else M.CYCLIC
    yb = M.CYCLIC if kind == "cyclic" else M.OPEN
    bcs = [(xb, xb), (yb, yb)]
    psi = W.random_field(grid, seed)
    if xb == M.CYCLIC:
        psi[-1] = psi[0]
    if yb == M.CYCLIC:
        psi[:, -1] = psi[:, 0]
    gc = W.streamfunction_advector(grid, seed + 100, periodic=(xb == M.CYCLIC, yb == M.CYCLIC),
                                   max_courant=0.45, exact_period=True)
    return bcs, psi, gc


def run(grid, opts, iters, dtype, bcs, psi, gc, steps, monkeypatch, pass2d):
    if pass2d:
        monkeypatch.delenv("MPDX_NO_PASS2D", raising=False)
    else:
        monkeypatch.setenv("MPDX_NO_PASS2D", "1")
    s = M.Solver(grid, opts=opts, n_iters=iters, dtype=dtype, bcs=bcs, max_abs_div_eps=1e300)
    s.set_advectee(psi)
    for d in range(2):
        s.set_advector(gc[d], d)
    s.advance(steps)
    mode = s.z_tiles
    out = s.advectee()
    s.close()
    return out, mode


def oracle_run(grid, opts, iters, bcs, psi, gc, steps):
    oo = (O.FCT if opts & M.FCT else 0) | (O.IGA if opts & M.IGA else 0)
    ob = [(O.CYCLIC if b[0] == M.CYCLIC else O.OPEN,) * 2 for b in bcs]
    o = O.Oracle(grid, opts=oo, n_iters=iters, bcs=ob, max_abs_div_eps=1e300)
    o.set_advectee(psi)
    for d in range(2):
        o.set_advector(gc[d], d)
    o.advance(steps)
    return o.advectee()


@pytest.mark.parametrize("kind", BOUNDS)
@pytest.mark.parametrize("grid", SIZES)
@pytest.mark.parametrize("opts,iters", OPTS)
def test_pass2d_matches_oracle_fp64(grid, opts, iters, kind, monkeypatch):
    bcs, psi, gc = case2d(grid, 100 + grid[0] + grid[1], kind)
    print(f"pass2d vs oracle grid={grid} opts={opts} n_iters={iters} bounds={kind}", flush=True)
    got, mode = run(grid, opts, iters, np.float64, bcs, psi, gc, 4, monkeypatch, True)
    assert mode == (3 if kind == "cyclic" else 0)
    want = oracle_run(grid, opts, iters, bcs, psi, gc, 4)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("kind", BOUNDS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("opts,iters", OPTS)
def test_pass2d_rounds_like_shell_kernels(opts, iters, dtype, kind, monkeypatch):
    grid = (64, 33)
    bcs, psi, gc = case2d(grid, 300, kind)
    print(f"pass2d vs shell kernels opts={opts} n_iters={iters} dtype={np.dtype(dtype).name} bounds={kind}",
          flush=True)
    a, _ = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, True)
    b, mb = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, False)
    assert mb == 0
    eps = np.finfo(dtype).eps
    assert np.abs(a.astype(np.float64) - b).max() <= 16 * eps * np.abs(b).max()


def test_pass2d_non_repeating_cyclic_fields_take_the_shell(monkeypatch):
    """Cyclic bounds but psi differing between planes 0 and n-1: the kernel
    covers the interior cells next to the shell, and the trajectory stays the
    oracle's."""
    grid = (40, 36)
    bcs, psi, gc = case2d(grid, 7, "cyclic")
    psi[-1, 5] += 0.25
    print("pass2d: cyclic, non-repeating psi", flush=True)
    got, mode = run(grid, M.FCT, 3, np.float64, bcs, psi, gc, 3, monkeypatch, True)
    assert mode == 0
    want = oracle_run(grid, M.FCT, 3, bcs, psi, gc, 3)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pass2d_deform_generator(dtype, monkeypatch):
    """The on-device deformational generator on an all-cyclic 2D grid: its
    corner tables repeat across the period, so the repeating psi keeps the
    pass kernel covering every cell."""
    grid = (60, 50)
    bcs = [(M.CYCLIC, M.CYCLIC)] * 2
    psi = W.random_field(grid, 900)
    psi[-1] = psi[0]
    psi[:, -1] = psi[:, 0]
    T, c, nt = 16.0, 0.45, 6
    print(f"pass2d deform generator dtype={np.dtype(dtype).name}", flush=True)
    monkeypatch.delenv("MPDX_NO_PASS2D", raising=False)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=[(O.CYCLIC, O.CYCLIC)] * 2, dt=1.0, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(2):
            x.set_advector(np.zeros(x.advector_shape(d)), d)
    o.set_generator(1, T, c)
    s.set_advector_generator(M.GC_DEFORM, T, c)
    o.advance(nt)
    s.advance(nt)
    assert s.z_tiles == 3
    got = s.advectee().astype(np.float64)
    s.close()
    want = o.advectee()
    assert np.abs(got - want).max() <= (1e-12 if dtype == np.float64 else 1e-5) * np.abs(want).max()


def test_pass2d_two_equations_hint_scale_fp64(monkeypatch):
    """Two equations with their own hint_scale (the kernel's scaled psi^n
    loads and the output scale) against the oracle."""
    monkeypatch.delenv("MPDX_NO_PASS2D", raising=False)
    grid = (48, 40)
    bcs, psi, gc = case2d(grid, 55, "cyclic")
    psi2 = psi[::-1].copy()
    print("pass2d two equations hint_scale", flush=True)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, bcs=bcs, n_eqns=2, dtype=np.float64, hint_scale=[3, -2],
                 max_abs_div_eps=1e300)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=[(O.CYCLIC, O.CYCLIC)] * 2, n_eqns=2, hint_scale=[3, -2],
                 max_abs_div_eps=1e300)
    for x in (s, o):
        x.set_advectee(psi, 0)
        x.set_advectee(psi2, 1)
        for d in range(2):
            x.set_advector(gc[d], d)
    s.advance(4)
    o.advance(4)
    assert s.z_tiles == 3
    for e in range(2):
        got, want = s.advectee(e), o.advectee(e)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    s.close()

"""Cyclic bounds in the interior kernel's tiles ("seam" modes).

With cyclic bounds on both z sides and fields that repeat plane z = 0 at
z = n-1 (the duplicated edge of a period of n-1 cells), the tiles also compute
the cells next to the z bounds, reading every halo column at its periodic
image (the value the cyclic halo fill puts there, bcond/cyclic_3d.hpp:29-47,
120-138), and the boundary shell has no z slabs (z_tiles 2).  With cyclic
bounds everywhere and fields repeating across every period, the tiles do the
same along x and y and compute every cell: no boundary shell, no halo fills
(z_tiles 3).  advance() compares those planes on the device at the start of
each call and takes these paths only when they are equal bit for bit (psi of
every equation and, without the on-device generator, the advector's halo
against its images).

In FP64 the results must match the oracle for every option set of the
interior kernel and z extents where one tile holds both bounds (12, 27, 28) or
the last tile overlaps its neighbour (29, 45, 61), with cyclic or open x and y
bounds.  Against the slab path (MPDX_NO_SEAM=1) they may differ only by
rounding: the interior and shell kernels compute a cell with the same
formulae, contracted into fused multiply-adds differently.  Fields that do not
repeat take the slab path."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu

OPTS = [(M.FCT, 2), (M.FCT, 3), (M.FCT, 4), (0, 3), (M.IGA, 2), (M.IGA | M.FCT, 2)]
NZ = [12, 27, 28, 29, 45, 61]


def consistent_case(grid, seed, xb, yb):
    """psi and a divergence-free advector repeating plane 0 at n-1 along z
    and along each cyclic one of x and y.
    Cyclic x and y: a 3D streamfunction flow (w != 0).  An open x or y
    bound: horizontal flow (w = 0), since the reference's zero-divergence
    halo faces at an open x or y edge are computed before the cyclic z fill
    has set the edge's z-halo faces of w (solver_3d.hpp:73-103 fills x, y,
    then z), which makes them differ between planes z = 0 and z = n-1
    unless w vanishes there; such fields take the slab path."""
    bcs = [(xb, xb), (yb, yb), (M.CYCLIC, M.CYCLIC)]
    psi = W.random_field(grid, seed)
    psi[:, :, -1] = psi[:, :, 0]
    if xb == M.CYCLIC:
        psi[-1] = psi[0]
    if yb == M.CYCLIC:
        psi[:, -1] = psi[:, 0]
    if xb == M.CYCLIC and yb == M.CYCLIC:
        gc = W.streamfunction_advector(grid, seed + 100, periodic=(True, True, True), max_courant=0.45,
                                       exact_period=True)
    else:
        gc = W.level_streamfunction_advector(grid, seed + 100, max_courant=0.45)
    for d in range(2):
        assert np.array_equal(gc[d][:, :, 0], gc[d][:, :, -1])
    return bcs, psi, gc


def expected_mode(bcs):
    return 3 if bcs[0][0] == M.CYCLIC and bcs[1][0] == M.CYCLIC else 2


def case(nz, seed):
    grid = (22, 19, nz)
    xb = M.CYCLIC if nz % 2 else M.OPEN
    yb = M.CYCLIC if nz % 3 == 0 and xb == M.CYCLIC else M.OPEN
    return (grid,) + consistent_case(grid, seed + nz, xb, yb)


def run(grid, opts, iters, dtype, bcs, psi, gc, steps, monkeypatch, seam, n_eqns=1, **kw):
    # (the corner-streamfunction advector is divergence-free in the domain,
    # not in every open-boundary halo face the reference's check covers)
    kw.setdefault("max_abs_div_eps", 1e300)
    if seam:
        monkeypatch.delenv("MPDX_NO_SEAM", raising=False)
    else:
        monkeypatch.setenv("MPDX_NO_SEAM", "1")
    s = M.Solver(grid, opts=opts, n_iters=iters, dtype=dtype, bcs=bcs, n_eqns=n_eqns, **kw)
    psis = psi if isinstance(psi, list) else [psi]
    for e, p in enumerate(psis):
        s.set_advectee(p, e)
    for d in range(3):
        s.set_advector(gc[d], d)
    s.advance(steps)
    mode = s.z_tiles
    out = [s.advectee(e) for e in range(n_eqns)]
    s.close()
    return (out if isinstance(psi, list) else out[0]), mode


def oracle_run(grid, opts, iters, bcs, psi, gc, steps, n_eqns=1, **kw):
    oo = (O.FCT if opts & M.FCT else 0) | (O.IGA if opts & M.IGA else 0)
    ob = [(O.CYCLIC if b[0] == M.CYCLIC else O.OPEN,) * 2 for b in bcs]
    kw.setdefault("max_abs_div_eps", 1e300)
    o = O.Oracle(grid, opts=oo, n_iters=iters, bcs=ob, n_eqns=n_eqns, **kw)
    psis = psi if isinstance(psi, list) else [psi]
    for e, p in enumerate(psis):
        o.set_advectee(p, e)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(steps)
    out = [o.advectee(e) for e in range(n_eqns)]
    return out if isinstance(psi, list) else out[0]


@pytest.mark.parametrize("nz", NZ)
@pytest.mark.parametrize("opts,iters", OPTS)
def test_seam_tiles_match_oracle_fp64(nz, opts, iters, monkeypatch):
    grid, bcs, psi, gc = case(nz, 100)
    print(f"seam vs oracle nz={nz} opts={opts} n_iters={iters} x/y bounds={bcs[0][0]}/{bcs[1][0]}", flush=True)
    got, mode = run(grid, opts, iters, np.float64, bcs, psi, gc, 4, monkeypatch, True)
    assert mode == expected_mode(bcs)
    want = oracle_run(grid, opts, iters, bcs, psi, gc, 4)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # the step keeps the planes equal
    assert np.array_equal(got[:, :, 0], got[:, :, -1])
    if mode == 3:
        assert np.array_equal(got[0], got[-1]) and np.array_equal(got[:, 0], got[:, -1])


@pytest.mark.parametrize("nz", NZ)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("opts,iters", OPTS)
def test_seam_tiles_round_like_slab_path(nz, dtype, opts, iters, monkeypatch):
    grid, bcs, psi, gc = case(nz, 300)
    print(f"seam vs slab nz={nz} dtype={np.dtype(dtype).name} opts={opts} n_iters={iters}", flush=True)
    a, ma = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, True)
    b, mb = run(grid, opts, iters, dtype, bcs, psi, gc, 4, monkeypatch, False)
    assert (ma, mb) == (expected_mode(bcs), 0)
    eps = np.finfo(dtype).eps
    assert np.abs(a.astype(np.float64) - b).max() <= 16 * eps * np.abs(b).max()


@pytest.mark.parametrize("yb", [M.OPEN, M.CYCLIC])
@pytest.mark.parametrize("iters", [3, 4])
def test_seam_tiles_several_tiles_and_chunks_fp64(iters, yb, monkeypatch):
    """A grid with several tiles along y and z and several x chunks (open
    y: the ring of tile rows next to the shell spans many tiles; cyclic y:
    the tiles wrap around in every dimension)."""
    grid = (70, 66, 93)
    bcs, psi, gc = consistent_case(grid, 17, M.CYCLIC, yb)
    print(f"seam several tiles grid={grid} n_iters={iters} y bounds={yb}", flush=True)
    got, mode = run(grid, M.FCT, iters, np.float64, bcs, psi, gc, 3, monkeypatch, True)
    assert mode == expected_mode(bcs)
    want = oracle_run(grid, M.FCT, iters, bcs, psi, gc, 3)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_seam_needs_repeating_planes(monkeypatch):
    """psi that differs between planes z = 0 and z = n-1 (or an advector
    whose planes differ) takes the slab path, which is what the reference
    computes for such input."""
    grid, bcs, psi, gc = case(29, 500)
    psi2 = psi.copy()
    psi2[3, 4, -1] += 0.25
    print("seam fallback: psi planes differ", flush=True)
    a, ma = run(grid, M.FCT, 3, np.float64, bcs, psi2, gc, 3, monkeypatch, True)
    b, mb = run(grid, M.FCT, 3, np.float64, bcs, psi2, gc, 3, monkeypatch, False)
    assert (ma, mb) == (0, 0)
    assert np.array_equal(a, b)
    gc2 = [g.copy() for g in gc]
    gc2[1][5, 6, -1] = gc2[1][5, 6, 0] + 2.0 ** -20
    print("seam fallback: advector planes differ", flush=True)
    _, mc = run(grid, M.FCT, 3, np.float64, bcs, psi, gc2, 1, monkeypatch, True)
    assert mc == 0
    # an open y bound with w != 0: the reference's open-edge halo faces
    # differ between the planes
    grid = (22, 19, 29)
    bcs = [(M.CYCLIC, M.CYCLIC), (M.OPEN, M.OPEN), (M.CYCLIC, M.CYCLIC)]
    gc3 = W.streamfunction_advector(grid, 9, periodic=(True, False, True), max_courant=0.45, exact_period=True)
    psi3 = W.random_field(grid, 10)
    psi3[:, :, -1] = psi3[:, :, 0]
    print("seam fallback: cyclic everywhere, psi repeating along z only", flush=True)
    grid4 = (24, 20, 27)
    bcs4, psi4, gc4 = consistent_case(grid4, 11, M.CYCLIC, M.CYCLIC)
    psi4[-1, 3, 5] += 0.125
    e4, me = run(grid4, M.FCT, 3, np.float64, bcs4, psi4, gc4, 3, monkeypatch, True)
    assert me == 2
    want4 = oracle_run(grid4, M.FCT, 3, bcs4, psi4, gc4, 3)
    assert np.abs(e4 - want4).max() <= 1e-12 * np.abs(want4).max()
    print("seam fallback: open y with w != 0", flush=True)
    d, md = run(grid, M.FCT, 3, np.float64, bcs, psi3, gc3, 3, monkeypatch, True)
    assert md == 0
    want = oracle_run(grid, M.FCT, 3, bcs, psi3, gc3, 3)
    assert np.abs(d - want).max() <= 1e-12 * np.abs(want).max()


def test_seam_decided_per_call(monkeypatch):
    """A call on repeating fields takes the seam tiles; after the caller
    breaks the repetition the next call takes the slab path, and the
    trajectory stays the oracle's."""
    grid, bcs, psi, gc = case(28, 700)
    print("seam decided per advance call", flush=True)
    monkeypatch.delenv("MPDX_NO_SEAM", raising=False)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, dtype=np.float64, bcs=bcs, max_abs_div_eps=1e300)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=[(O.CYCLIC if b[0] == M.CYCLIC else O.OPEN,) * 2 for b in bcs],
                 max_abs_div_eps=1e300)
    for x in (s, o):
        x.set_advectee(psi)
        for d in range(3):
            x.set_advector(gc[d], d)
    s.advance(2)
    o.advance(2)
    assert s.z_tiles == expected_mode(bcs)
    p = s.advectee()
    p[2, 3, 0] += 0.5
    s.set_advectee(p)
    o.set_advectee(p)
    s.advance(2)
    o.advance(2)
    assert s.z_tiles == 0
    got = s.advectee()
    s.close()
    want = o.advectee()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("planes", ["z", "xyz"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_seam_with_deform_generator(dtype, planes, monkeypatch):
    """The on-device generator (z-independent, its corner tables repeating
    across the x and y periods) keeps the advector's planes equal: cyclic in
    all dimensions (the multi-tracer workload's bounds), psi repeating along
    z only (z_tiles 2) or along every dimension (3)."""
    grid = (40, 34, 14)
    bcs = [(M.CYCLIC, M.CYCLIC)] * 3
    psi = W.random_field(grid, 801)
    psi[:, :, -1] = psi[:, :, 0]
    if planes == "xyz":
        psi[-1] = psi[0]
        psi[:, -1] = psi[:, 0]
    T, c, nt = 16.0, 0.45, 6
    print(f"seam deform generator dtype={np.dtype(dtype).name} planes={planes}", flush=True)
    monkeypatch.delenv("MPDX_NO_SEAM", raising=False)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=[(O.CYCLIC, O.CYCLIC)] * 3, dt=1.0, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(3):
            x.set_advector(np.zeros(x.advector_shape(d)), d)
    o.set_generator(1, T, c)
    s.set_advector_generator(M.GC_DEFORM, T, c)
    o.advance(nt)
    s.advance(nt)
    assert s.z_tiles == (3 if planes == "xyz" else 2)
    got = s.advectee().astype(np.float64)
    s.close()
    want = o.advectee()
    assert np.abs(got - want).max() <= (1e-12 if dtype == np.float64 else 1e-5) * np.abs(want).max()


@pytest.mark.parametrize("batched", [False, True])
def test_seam_two_equations_hint_scale_fp64(batched, monkeypatch):
    """Two equations with their own hint_scale, stepped one after the other
    or as batched launches (MPDX_BATCH_EQNS=1), against the oracle."""
    grid = (24, 20, 33)
    bcs, psi, gc = consistent_case(grid, 41, M.CYCLIC, M.CYCLIC)
    psi2 = psi[::-1, :, :].copy()
    print(f"seam two equations batched={batched}", flush=True)
    if batched:
        monkeypatch.setenv("MPDX_BATCH_EQNS", "1")
    else:
        monkeypatch.delenv("MPDX_BATCH_EQNS", raising=False)
    got, mode = run(grid, M.FCT, 4, np.float64, bcs, [psi, psi2], gc, 4, monkeypatch, True, n_eqns=2,
                    hint_scale=[2, -1])
    assert mode == 3
    want = oracle_run(grid, M.FCT, 4, bcs, [psi, psi2], gc, 4, n_eqns=2, hint_scale=[2, -1])
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= 1e-12 * np.abs(w).max()


def test_seam_ensemble_member_matches_solo(monkeypatch):
    """An ensemble member on repeating cyclic-z fields steps through the
    seam tiles inside the batched launches, bit for bit like its own
    advance()."""
    monkeypatch.delenv("MPDX_NO_SEAM", raising=False)
    grid = (30, 28, 40)
    bcs, psi, gc = consistent_case(grid, 61, M.CYCLIC, M.CYCLIC)
    grid2 = (26, 24, 20)
    bcs2, psi2, gc2 = consistent_case(grid2, 62, M.OPEN, M.CYCLIC)
    print("seam ensemble member vs solo", flush=True)

    def make(g, b, p, u):
        s = M.Solver(g, opts=M.FCT, n_iters=3, dtype=np.float32, bcs=b, max_abs_div_eps=1e300)
        s.set_advectee(p)
        for d in range(3):
            s.set_advector(u[d], d)
        return s

    solo = []
    for g, b, p, u in ((grid, bcs, psi, gc), (grid2, bcs2, psi2, gc2)):
        s = make(g, b, p, u)
        s.advance(5)
        assert s.z_tiles == expected_mode(b)
        solo.append(s.advectee())
        s.close()
    members = [make(grid, bcs, psi, gc), make(grid2, bcs2, psi2, gc2)]
    ens = M.Ensemble(members)
    ens.advance(5)
    for s, b, want in zip(members, (bcs, bcs2), solo):
        assert s.z_tiles == expected_mode(b)
        assert np.array_equal(s.advectee(), want)
    ens.close()
    for s in members:
        s.close()

"""GPU parity: the HIP library against the FP64 CPU oracle (same inputs).

FP64: agreement to <= 1e-12 relative to the field's max (the oracle and the
GPU differ only by fused multiply-add contraction and division rounding).
FP32: L2 and Linf against the FP64 oracle after a few steps.
Inputs: seeded positive random fields (not periodic-consistent, so every
halo replay rule is exercised) and streamfunction advectors with exact
zero divergence.
"""
import itertools
import sys

import numpy as np
import pytest

import oracle as O
from gpu_util import rel_err, run_gpu, run_oracle, setup_case
from libmpdataxx_amd.workloads import random_field as W_random

pytestmark = pytest.mark.gpu

OPTS = {"upwind": (0, 1), "basic": (0, 2), "iters3": (0, 3), "fct": (O.FCT, 2),
        "fct3": (O.FCT, 3), "iga": (O.IGA, 2), "iga_fct": (O.IGA | O.FCT, 2),
        "abs": (O.ABS, 2), "abs_fct3": (O.ABS | O.FCT, 3), "npa_pfc_fct": (O.NPA | O.PFC | O.FCT, 3),
        "khn_fct": (O.KHN | O.FCT, 2)}
BCS2 = [[(0, 0), (0, 0)], [(1, 1), (1, 1)], [(0, 0), (1, 1)], [(1, 1), (0, 0)]]
BCS3 = [[(0, 0)] * 3, [(1, 1)] * 3, [(1, 1), (0, 0), (1, 1)], [(0, 0), (1, 1), (0, 0)]]


def _case(nd, bcs, name, seed):
    opts, iters = OPTS[name]
    grid = (13, 11) if nd == 2 else (9, 8, 10)
    psi, gc = setup_case(grid, bcs, seed)
    print(f"case nd={nd} bcs={bcs} opts={name} seed={seed}", flush=True)
    return grid, opts, iters, psi, gc


@pytest.mark.parametrize("name", list(OPTS))
@pytest.mark.parametrize("bi", range(4))
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("path", [1, 2])
def test_fp64_matches_oracle(path, nd, bi, name):
    """path 1 = per-pass kernels, path 2 = fused kernels (abs, npa, pfc and khn
    on the boundary-shell kernels over the whole grid)."""
    bcs = (BCS2 if nd == 2 else BCS3)[bi]
    grid, opts, iters, psi, gc = _case(nd, bcs, name, 10 * bi + nd)
    want = run_oracle(grid, bcs, opts, iters, psi, gc, 4)[0]
    got = run_gpu(grid, bcs, opts, iters, psi, gc, 4, path=path)[0]
    assert not np.isnan(got).any()
    assert rel_err(got, want) <= 1e-12


@pytest.mark.parametrize("bi", [0, 1])
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("path", [1, 2])
def test_khn_donor_bit_exact(path, nd, bi):
    """khn with the donor-cell pass alone (n_iters = 1): no division occurs
    and no product can be contracted differently (one of pospart / negpart of
    each face velocity is zero), so the GPU's compensated sums equal the
    oracle's (pinned against math.fsum) bit for bit.  Control: the same run
    without khn differs from them (psi spans eight decades, so the pairwise
    sum rounds differently in some cells)."""
    bcs = (BCS2 if nd == 2 else BCS3)[bi]
    grid = (40, 36) if nd == 2 else (20, 18, 22)
    rng = np.random.default_rng(17 + 2 * nd + bi)
    psi = 10.0 ** rng.uniform(-4.0, 4.0, grid)
    _, gc = setup_case(grid, bcs, 3 + bi)
    print(f"khn donor nd={nd} bcs={bcs} path={path}", flush=True)
    want = run_oracle(grid, bcs, O.KHN, 1, psi, gc, 3)[0]
    got = run_gpu(grid, bcs, O.KHN, 1, psi, gc, 3, path=path)[0]
    assert np.array_equal(got, want)
    plain = run_gpu(grid, bcs, 0, 1, psi, gc, 3, path=path)[0]
    assert not np.array_equal(plain, want)


@pytest.mark.parametrize("name", ["basic", "fct", "fct3", "iga_fct"])
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("path", [1, 2])
def test_fp32_close_to_oracle(path, nd, name):
    bcs = (BCS2 if nd == 2 else BCS3)[1]
    grid, opts, iters, psi, gc = _case(nd, bcs, name, 77)
    want = run_oracle(grid, bcs, opts, iters, psi, gc, 4)[0]
    got = run_gpu(grid, bcs, opts, iters, psi, gc, 4, dtype=np.float32, path=path)[0]
    assert rel_err(got, want) <= 1e-5


@pytest.mark.parametrize("path", [1, 2])
def test_multi_equation_and_hint_scale(path):
    grid, bcs = (9, 8, 10), [(0, 0)] * 3
    psi, gc = setup_case(grid, bcs, 5)
    psi2 = psi[::-1].copy()
    want = run_oracle(grid, bcs, O.FCT, 3, [psi, psi2], gc, 3, n_eqns=2, hint_scale=[0, 3])
    got = run_gpu(grid, bcs, O.FCT, 3, [psi, psi2], gc, 3, n_eqns=2, hint_scale=[0, 3], path=path)
    for g, w in zip(got, want):
        assert rel_err(g, w) <= 1e-12


@pytest.mark.parametrize("path", [1, 2])
def test_upwind_filter_freq(path):
    grid, bcs = (13, 11), [(1, 1)] * 2
    psi, gc = setup_case(grid, bcs, 9)
    want = run_oracle(grid, bcs, O.FCT, 3, psi, gc, 5, upwind_filter_freq=2)[0]
    got = run_gpu(grid, bcs, O.FCT, 3, psi, gc, 5, upwind_filter_freq=2, path=path)[0]
    assert rel_err(got, want) <= 1e-12


@pytest.mark.parametrize("nd", [2, 3])
def test_fused_larger_grid_fp64(nd):
    """A grid spanning several kernel bricks in every direction."""
    grid = (70, 45) if nd == 2 else (21, 19, 70)
    bcs = [(1, 1)] * nd
    psi, gc = setup_case(grid, bcs, 31)
    want = run_oracle(grid, bcs, O.FCT, 3, psi, gc, 3)[0]
    got = run_gpu(grid, bcs, O.FCT, 3, psi, gc, 3, path=2)[0]
    assert rel_err(got, want) <= 1e-12


INTERIOR_OPTS = ["basic", "iters3", "fct", "fct3", "iga", "iga_fct"]


# (element type, guard mode): FP64, FP32 with the default FP64-constant
# guards, FP32 with the guarded ratios evaluated in FP64
PRECISIONS = {"f64": (np.float64, 0), "f32": (np.float32, 0), "f32_exact": (np.float32, 2)}


@pytest.mark.parametrize("name", INTERIOR_OPTS)
@pytest.mark.parametrize("bi", range(4))
@pytest.mark.parametrize("prec", list(PRECISIONS))
def test_fused_interior_matches_oracle(prec, bi, name):
    """3D grids large enough for the x-marching interior kernel, with several
    x chunks and y/z tiles, every boundary combination; FP64 within 1e-12 of
    the oracle, FP32 (either guard mode) within 1e-5."""
    dtype, guard = PRECISIONS[prec]
    opts, iters = OPTS[name]
    bcs = BCS3[bi]
    grid = [(140, 14, 36), (14, 70, 36), (30, 26, 75), (40, 33, 20)][bi]
    psi, gc = setup_case(grid, bcs, 100 + bi)
    print(f"interior case grid={grid} bcs={bcs} opts={name} prec={prec}", flush=True)
    want = run_oracle(grid, bcs, opts, iters, psi, gc, 3)[0]
    got = run_gpu(grid, bcs, opts, iters, psi, gc, 3, dtype=dtype, path=2, guard=guard)[0]
    assert not np.isnan(got).any()
    assert rel_err(got, want) <= (1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_graph_replay_matches_eager(dtype, monkeypatch):
    """Time steps replayed from the captured hipGraphs (one per time-level
    parity) give bit-identical fields to eagerly launched steps, over several
    replays of each graph with host reads, reductions and re-uploads of the
    advectee between them; running the boundary shell on the auxiliary stream
    or on the main stream gives the same fields."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 75), BCS3[2]
    psi, gc = setup_case(grid, bcs, 123)
    psi2 = psi[:, ::-1].copy()
    print(f"graph replay grid={grid} dtype={np.dtype(dtype).name}", flush=True)

    def make(no_graphs):
        if no_graphs:
            monkeypatch.setenv("MPDX_NO_GRAPHS", "1")
        else:
            monkeypatch.delenv("MPDX_NO_GRAPHS", raising=False)
        s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, n_eqns=2, dtype=dtype,
                     hint_scale=[0, 2], max_abs_div_eps=1e300, path=2)
        s.set_advectee(psi, 0)
        s.set_advectee(psi2, 1)
        for d in range(3):
            s.set_advector(gc[d], d)
        return s

    eager = make(True)
    eager.advance(7)
    want = [eager.advectee(e) for e in range(2)]
    eager.close()
    # the boundary shell on the auxiliary stream or on the main stream
    monkeypatch.setenv("MPDX_NO_CONCURRENCY", "1")
    serial = make(True)
    serial.advance(7)
    for e in range(2):
        assert np.array_equal(serial.advectee(e), want[e])
    serial.close()
    monkeypatch.delenv("MPDX_NO_CONCURRENCY")

    g = make(False)
    masses = []
    for step in range(7):
        g.advance(1)
        masses.append(g.mass(0))
        if step == 3:
            # a host round trip of the advectee between replays
            g.set_advectee(g.advectee(1), 1)
    got = [g.advectee(e) for e in range(2)]
    g.close()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert all(np.isfinite(masses))


DEFORM_CASES = [((40, 34, 12), [(0, 0)] * 3), ((26, 22, 14), [(1, 1), (0, 0), (1, 1)]),
                ((60, 50), [(0, 0)] * 2)]


@pytest.mark.parametrize("case", range(len(DEFORM_CASES)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deform_generator_matches_oracle(dtype, case):
    """The on-device reversing deformational advector (recomputed at every
    step, then halo-refilled like the reference's calc_gc) against the
    oracle's generator: same corner values, same trajectory."""
    import libmpdataxx_amd as M
    grid, bcs = DEFORM_CASES[case]
    psi = W_random(grid, 300 + case)
    T, c, nt = 16.0, 0.45, 6
    print(f"deform generator grid={grid} bcs={bcs} dtype={np.dtype(dtype).name}", flush=True)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(len(grid)):
            x.set_advector(np.zeros(x.advector_shape(d)), d)
    o.set_generator(1, T, c)
    s.set_advector_generator(M.GC_DEFORM, T, c)
    o.advance(nt)
    s.advance(nt)
    for d in range(len(grid)):
        assert np.array_equal(s.advector(d).astype(np.float64), o.advector(d))
    got = s.advectee().astype(np.float64)
    s.close()
    assert rel_err(got, o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_calc_gc_callback_matches_oracle(dtype):
    """A host calc_gc callback that uploads a new advector every step: the
    library refills the advector halos like the reference's calc_gc, so the
    trajectory matches the oracle's generated-advector run."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 16), [(0, 0), (1, 1), (0, 0)]
    psi = W_random(grid, 401)
    T, c, nt = 12.0, 0.4, 5
    print(f"calc_gc callback grid={grid} dtype={np.dtype(dtype).name}", flush=True)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, max_abs_div_eps=1e300)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(np.zeros(o.advector_shape(d)), d)
    o.set_generator(1, T, c)
    gcs = []
    for _ in range(nt):
        o.advance(1)
        gcs.append([o.advector(d) for d in range(3)])

    s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(np.zeros(s.advector_shape(d)), d)
    calls = []

    def calc_gc(t_half):
        n = int(round(t_half - 0.5))
        calls.append(t_half)
        for d in range(3):
            s.set_advector(gcs[n][d], d)
        return True

    s.set_calc_gc(calc_gc)
    s.advance(2)
    s.advance(nt - 2)
    got = s.advectee().astype(np.float64)
    s.close()
    assert calls == [k + 0.5 for k in range(nt)]
    assert rel_err(got, o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_var_dt_matches_oracle(dtype, path):
    """var_dt: the initial rescaling of dt and the advector to max_courant,
    then steps until the requested duration, a duration of 0 and a step
    count, against the oracle."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 20), [(1, 1), (0, 0), (1, 1)]
    psi, gc = setup_case(grid, bcs, 55)
    print(f"var_dt dtype={np.dtype(dtype).name} path={path}", flush=True)
    kw = dict(opts=O.FCT, n_iters=3, bcs=bcs, dt=0.2, max_abs_div_eps=1e300, var_dt=True, max_courant=0.42)
    o = O.Oracle(grid, **kw)
    s = M.Solver(grid, dtype=dtype, path=path, **kw)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(3):
            x.set_advector(gc[d], d)
    o.advance_time(1.5)
    s.advance_time(1.5)
    assert s.timestep == o.timestep
    assert abs(s.dt - o.dt) <= 1e-15 * o.dt * (1 if dtype == np.float64 else 1e8)
    assert rel_err(s.advectee().astype(np.float64), o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)
    # a duration of 0 takes one step (solver_common.hpp:293-294,352); a step
    # count takes that many steps of the current dt
    ts = o.timestep
    for x in (o, s):
        x.advance_time(0.0)
        x.advance(2)
    assert s.timestep == o.timestep == ts + 3
    assert abs(s.time - o.time) <= 1e-12 * o.time
    assert rel_err(s.advectee().astype(np.float64), o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)
    s.close()


@pytest.mark.parametrize("ord_", [1, 2])
@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_var_dt_output_matches_oracle(dtype, path, ord_):
    """var_dt output (output_common.hpp:47-62,72-140): the start call and
    every record (kind, time, timestep, interpolated field of both
    equations) of the library equal the oracle's; FP64 <= 1e-12, FP32 <= 1e-5
    relative.  Two advance calls (durations 1.3 and 0.9, plus the further
    step of out_intrp_ord 2), records every 0.5."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 20), [(1, 1), (0, 0), (1, 1)]
    psi, gc = setup_case(grid, bcs, 56)
    psi1 = W_random(grid, 57)
    print(f"var_dt output dtype={np.dtype(dtype).name} path={path} ord={ord_}", flush=True)
    kw = dict(opts=O.FCT, n_iters=3, n_eqns=2, bcs=bcs, dt=0.2, max_abs_div_eps=1e300, var_dt=True,
              max_courant=0.42)
    o = O.Oracle(grid, **kw)
    s = M.Solver(grid, dtype=dtype, path=path, **kw)
    recs = {id(o): [], id(s): []}
    for x in (o, s):
        x.set_advectee(psi, 0)
        x.set_advectee(psi1, 1)
        for d in range(3):
            x.set_advector(gc[d], d)

        def out(kind, t, x=x):
            recs[id(x)].append((kind, t, x.timestep, [x.output(e).astype(np.float64) for e in (1, 0)]))
        x.set_var_dt_output(out, 0.5, ord_, outvars=(1, 0))
    for dur in (1.3, 0.9):
        for x in (o, s):
            x.advance_time(dur)
            if ord_ == 2:
                x.advance(1)
    ro, rs = recs[id(o)], recs[id(s)]
    assert [r[0] for r in ro] == [0] + [1] * (len(ro) - 1) and len(ro) >= 4
    assert [r[:3] for r in rs] == [r[:3] for r in ro]
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for a, b in zip(rs, ro):
        for fa, fb in zip(a[3], b[3]):
            assert rel_err(fa, fb) <= tol, (a[:3], rel_err(fa, fb))
    s.close()


def test_var_dt_with_calc_gc_rescales_every_step():
    """var_dt with a time-varying advector: the callback writes velocity x
    dt; the library shrinks dt until the Courant number is max_courant
    (solver_common.hpp:310-324), and the run equals a fixed-dt run at that
    dt."""
    import libmpdataxx_amd as M
    grid, bcs = (24, 20, 18), [(0, 0)] * 3
    psi, vel = setup_case(grid, bcs, 56)     # vel: Courant numbers per unit dt
    print("var_dt with calc_gc", flush=True)
    s = M.Solver(grid, opts=M.FCT, n_iters=2, bcs=bcs, dt=1.0, var_dt=True, max_courant=0.3,
                 max_abs_div_eps=1e300)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(vel[d], d)

    def calc_gc(t_half):
        for d in range(3):
            s.set_advector(vel[d] * s.dt, d)
        return True
    s.set_calc_gc(calc_gc)
    s.advance_time(10.0)
    dt = s.dt
    steps = s.timestep
    got = s.advectee()
    s.close()
    f = M.Solver(grid, opts=M.FCT, n_iters=2, bcs=bcs, dt=dt, max_abs_div_eps=1e300)
    f.set_advectee(psi)
    for d in range(3):
        f.set_advector(vel[d] * dt, d)
    f.advance(steps)
    assert rel_err(f.advectee(), got) <= 1e-12
    f.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_generator_across_growing_advance_calls(dtype):
    """advance(2) then advance(10) with the on-device generator: the second
    call needs a longer amplitude table than the first captured its graph
    with, so the graphs must be recaptured; the advector and trajectory equal
    the oracle's 12 steps (advector bit for bit)."""
    import libmpdataxx_amd as M
    grid, bcs = (40, 34, 14), [(0, 0)] * 3
    psi = W_random(grid, 310)
    T, c = 20.0, 0.4
    print(f"generator advance(2)+advance(10) dtype={np.dtype(dtype).name}", flush=True)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(3):
            x.set_advector(np.zeros(x.advector_shape(d)), d)
    o.set_generator(1, T, c)
    s.set_advector_generator(M.GC_DEFORM, T, c)
    o.advance(12)
    s.advance(2)
    s.advance(10)
    for d in range(3):
        assert np.array_equal(s.advector(d).astype(np.float64), o.advector(d))
    got = s.advectee().astype(np.float64)
    s.close()
    assert rel_err(got, o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_generator_switched_on_after_first_advance(dtype):
    """Steps with a fixed advector (captured graphs without the generator),
    then the generator switched on: the next steps must run it."""
    import libmpdataxx_amd as M
    grid, bcs = (40, 34, 14), [(0, 0)] * 3
    psi, gc = setup_case(grid, bcs, 311)
    T, c = 20.0, 0.4
    print(f"generator switched on dtype={np.dtype(dtype).name}", flush=True)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, dt=1.0, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi)
        for d in range(3):
            x.set_advector(gc[d], d)
    o.advance(3)
    s.advance(3)
    o.set_generator(1, T, c)
    s.set_advector_generator(M.GC_DEFORM, T, c)
    o.advance(4)
    s.advance(4)
    for d in range(3):
        assert np.array_equal(s.advector(d).astype(np.float64), o.advector(d))
    got = s.advectee().astype(np.float64)
    s.close()
    assert rel_err(got, o.advectee()) <= (1e-12 if dtype == np.float64 else 1e-5)


def test_var_dt_calc_gc_unchanged_keeps_dt():
    """A calc_gc hook that reports an unchanged advector: the reference
    rescales dt only when calc_gc returns true (solver_common.hpp:306-324),
    so dt, time and the field equal a var_dt run without the hook."""
    import libmpdataxx_amd as M
    grid, bcs = (24, 20, 18), [(0, 0)] * 3
    psi, gc = setup_case(grid, bcs, 57)
    print("var_dt with an unchanged calc_gc", flush=True)
    runs = []
    for hook in (False, True):
        s = M.Solver(grid, opts=M.FCT, n_iters=2, bcs=bcs, dt=0.5, var_dt=True, max_courant=0.3,
                     max_abs_div_eps=1e300)
        s.set_advectee(psi)
        for d in range(3):
            s.set_advector(gc[d], d)
        calls = []
        if hook:
            s.set_calc_gc(lambda t_half: calls.append(t_half) and False)
        s.advance_time(4.0)
        runs.append((s.dt, s.time, s.timestep, s.advectee(), len(calls)))
        s.close()
    (dt0, t0, n0, a0, _), (dt1, t1, n1, a1, ncalls) = runs
    assert ncalls == n1 > 0
    assert dt1 == dt0 and t1 == t0 and n1 == n0
    assert np.array_equal(a1, a0)


def test_batched_equations_match_sequential(monkeypatch):
    """MPDX_BATCH_EQNS=1: the captured step records each equation's launches
    and merges them into batched launches (each equation with velocity
    scratch of its own); the fields equal the default sequential steps bit
    for bit, over several replays of both graphs."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 40), BCS3[0]
    psi, gc = setup_case(grid, bcs, 124)
    fields = [psi, psi[::-1].copy(), psi[:, ::-1].copy()]
    print("batched equations vs sequential", flush=True)
    outs = []
    for batched in (False, True):
        if batched:
            monkeypatch.setenv("MPDX_BATCH_EQNS", "1")
        s = M.Solver(grid, opts=O.FCT, n_iters=3, bcs=bcs, n_eqns=3, dtype=np.float32, hint_scale=[0, 2, -1],
                     max_abs_div_eps=1e300, path=2)
        for e, f in enumerate(fields):
            s.set_advectee(f, e)
        for d in range(3):
            s.set_advector(gc[d], d)
        for _ in range(5):
            s.advance(1)
        outs.append([s.advectee(e) for e in range(3)])
        s.close()
        monkeypatch.delenv("MPDX_BATCH_EQNS", raising=False)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)

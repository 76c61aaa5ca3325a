"""The reference's own FCT unit tests on the GPU paths (inputs in
tests/reference_units.py): cone_bugs (tests/unit/cone_bugs/test_issue.cpp,
the minimum stays exactly 1 after 2 steps) and the bounds of hint_scale
(tests/unit/hint_scale/hint_scale.cpp, hint_scale = 24: max <= the initial
max, min >= 0), in FP32 with the reference's float semantics and with the
default guards, and in FP64 against the oracle; and the convergence orders
of tests/sandbox/convergence_2d_3d (2D and 3D); the order independence of
the compensated mass sum (tests/unit/kahan_sum); and the orders and the
variable-time-step gain of tests/unit/var_dt."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from reference_units import (CONVERGENCE_SIZES, F32_EPS, cone_bugs_case, convergence_case,
                             hint_scale_case, l2_error)

pytestmark = pytest.mark.gpu

VARIANTS = [(np.float32, M.GUARD_NATIVE), (np.float32, M.GUARD_F64), (np.float32, M.GUARD_F64_EXACT),
            (np.float64, M.GUARD_F64)]
PATHS = [M.PATH_FUSED, M.PATH_PERPASS]


def run_gpu(case, opts, dtype, guard, path, hint=0):
    s = M.Solver(case["grid"], opts=opts, n_iters=case["n_iters"], bcs=case["bcs"], dtype=dtype,
                 guard=guard, path=path, hint_scale=[hint], max_abs_div_eps=F32_EPS)
    s.set_advectee(case["psi"])
    for d in range(len(case["grid"])):
        s.set_advector(case["gc"][d], d)
    s.advance(case["nt"])
    out = s.advectee().astype(np.float64)
    s.close()
    return out


def run_oracle(case, opts, hint=0):
    o = O.Oracle(case["grid"], opts=opts, n_iters=case["n_iters"], bcs=case["bcs"],
                 hint_scale=[hint], max_abs_div_eps=F32_EPS)
    o.set_advectee(case["psi"])
    for d in range(len(case["grid"])):
        o.set_advector(case["gc"][d], d)
    o.advance(case["nt"])
    return o.advectee()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype,guard", VARIANTS)
def test_cone_bugs_minimum_stays_one(dtype, guard, path):
    c = cone_bugs_case()
    print(f"cone_bugs dtype={np.dtype(dtype).name} guard={guard} path={path}", flush=True)
    out = run_gpu(c, M.FCT, dtype, guard, path)
    assert out.min() == 1.0
    if dtype == np.float64:
        want = run_oracle(c, M.FCT)
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype,guard", VARIANTS)
def test_hint_scale_keeps_bounds(dtype, guard, path):
    c = hint_scale_case()
    print(f"hint_scale dtype={np.dtype(dtype).name} guard={guard} path={path}", flush=True)
    out = run_gpu(c, M.FCT | M.IGA, dtype, guard, path, hint=24)
    assert np.isfinite(out).all()
    assert out.max() <= c["psi"].max()
    assert out.min() >= 0.0
    if dtype == np.float64:
        want = run_oracle(c, M.FCT | M.IGA, hint=24)
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("opts", [0, M.IGA | M.FCT])
@pytest.mark.parametrize("nd", [2, 3])
def test_convergence_orders(nd, opts, dtype, path):
    print(f"convergence nd={nd} opts={opts} dtype={np.dtype(dtype).name} path={path}", flush=True)
    errs = []
    for np_ in CONVERGENCE_SIZES[nd]:
        c = convergence_case(np_, nd)
        s = M.Solver(c["grid"], opts=opts, n_iters=c["n_iters"], bcs=c["bcs"], dtype=dtype, path=path)
        s.set_advectee(c["psi"])
        for d in range(nd):
            s.set_advector(c["gc"][d], d)
        s.advance(c["nt"])
        errs.append(l2_error(s.advectee().astype(np.float64), c))
        s.close()
    assert round(np.log2(errs[0] / errs[1])) == 2, errs


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_kahan_mass_sum_order_independent(dtype):
    # tests/unit/kahan_sum/test_kahan_sum.cpp: one 1e8 and many 1e-8; the
    # compensated sums of the field and of its reverse agree within 1.4e-8
    # (an uncompensated sum loses the small terms in one order only)
    print(f"kahan mass sum dtype={np.dtype(dtype).name}", flush=True)
    grid = (10, 6, 5)
    a = np.full(grid, 1e-8, dtype=dtype)
    a[0] = 1e8
    sums = []
    for field in (a, a[::-1].copy()):
        s = M.Solver(grid, opts=M.FCT, n_iters=2, dtype=dtype)
        s.set_advectee(field.astype(np.float64))
        sums.append(s.reduce(M.RED_SUM))
        s.close()
    import math
    exact = math.fsum(a.astype(np.float64).ravel())
    assert abs(sums[0] - sums[1]) <= 1.4e-8
    assert abs(sums[0] - exact) <= np.spacing(exact)


def _var_dt_error(np_, cfl, var_dt):
    """tests/unit/var_dt/var_dt_test.cpp:18-68 in double, the 1D grid as
    np x 3 cyclic with no flow along y: psi = 2 + sin(i dx), dx = 2 pi /
    (np - 1), calc_gc sets GC_x = dt / dx u_0 cos(pi t_half / tau) (tau = 256,
    u_0 = 1) each step with the current dt; time = tau sqrt(2) / 2; relative
    L2 error against 2 + sin(i dx - sin(pi u_0 t / tau) tau / pi)."""
    import ctypes
    pi, tau, u0 = np.pi, 256.0, 1.0
    time = tau * np.sqrt(2.0) / 2
    dx = 2.0 * pi / (np_ - 1)
    dt = cfl * dx
    grid = (np_, 3)
    s = M.Solver(grid, opts=0, n_iters=2, bcs=[(M.CYCLIC, M.CYCLIC)] * 2, dtype=np.float64,
                 dt=dt, var_dt=var_dt, max_courant=cfl)
    x = np.arange(np_) * dx
    s.set_advectee(np.repeat((2.0 + np.sin(x))[:, None], 3, axis=1))
    s.set_advector(np.zeros(s.advector_shape(0)), 0)
    s.set_advector(np.zeros(s.advector_shape(1)), 1)
    cur = ctypes.c_double()
    L = M.lib()
    shape0 = s.advector_shape(0)

    def calc_gc(t_half):
        L.mpdx_dt(s._h, ctypes.byref(cur))
        s.set_advector(np.full(shape0, cur.value / dx * u0 * np.cos(pi * t_half / tau)), 0)
        return True

    s.set_calc_gc(calc_gc)
    if var_dt:
        s.advance_time(time)
    else:
        s.advance(int(time / dt))
    t = s.time
    got = s.advectee()[:, 1]
    s.close()
    true = 2 + np.sin(x - np.sin(pi * u0 * t / tau) * tau / pi)
    return np.sqrt(np.sum((true - got) ** 2) / np.sum(true ** 2))


@pytest.mark.parametrize("cfl", [0.1, 0.5, 0.9])
def test_var_dt_orders_and_gain(cfl):
    # var_dt_test.cpp:70-95: order 2 with a constant and with a variable
    # time step, and the variable one more accurate on the fine grid
    print(f"var_dt cfl={cfl}", flush=True)
    e = {(n, v): _var_dt_error(n, cfl, v) for n in (33, 65) for v in (False, True)}
    ord_cdt = round(np.log2(e[33, False] / e[65, False]))
    ord_vdt = round(np.log2(e[33, True] / e[65, True]))
    assert ord_cdt == 2 and ord_vdt == 2, e
    assert e[65, True] < e[65, False], e


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_basic_example_log(dtype, path, golden_dir):
    # paper_2015_GMD/0_basic_example: the six printed records (5 decimals)
    from reference_units import basic_example_case, basic_example_records
    print(f"basic_example dtype={np.dtype(dtype).name} path={path}", flush=True)
    c = basic_example_case()
    want = basic_example_records(golden_dir)
    s = M.Solver(c["grid"], opts=M.IGA | M.FCT, n_iters=c["n_iters"], bcs=c["bcs"], dtype=dtype, path=path)
    s.set_advectee(c["psi"])
    for d in range(2):
        s.set_advector(c["gc"][d], d)
    for k, rec in enumerate(want):
        if k:
            s.advance(c["outfreq"])
        got = s.advectee()[:, 1].astype(np.float64)
        assert np.abs(got - rec).max() <= 5e-6 + (1e-6 if dtype == np.float32 else 1e-12), k
    s.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("group,opts", [(0, M.ABS), (1, M.IGA), (3, M.IGA | M.FCT)])
def test_advscheme_opts_log(group, opts, dtype, path, golden_dir):
    # paper_2015_GMD/1_advscheme_opts: both equations after 1200 steps
    from reference_units import advscheme_opts_case, advscheme_opts_finals
    print(f"advscheme_opts group={group} dtype={np.dtype(dtype).name} path={path}", flush=True)
    c = advscheme_opts_case()
    want = advscheme_opts_finals(golden_dir, group)
    s = M.Solver(c["grid"], opts=opts, n_iters=c["n_iters"], n_eqns=2, bcs=c["bcs"], dtype=dtype, path=path)
    for e in range(2):
        s.set_advectee(c["psi"][e], e)
    for d in range(2):
        s.set_advector(c["gc"][d], d)
    s.advance(c["nt"])
    # FP64: the printed 5 decimals; FP32: its rounding accumulated over 1200
    # steps of the non-monotone iga scheme at a step (~2e-5 relative measured)
    tol = 5e-6 + (2e-4 if dtype == np.float32 else 1e-12)
    for e in range(2):
        assert np.abs(s.advectee(e)[:, 1].astype(np.float64) - want[e]).max() <= tol, e
    s.close()

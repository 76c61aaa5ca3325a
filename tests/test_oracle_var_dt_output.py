"""The oracle's var_dt output (output/detail/output_common.hpp:47-62,72-140)
pinned where the interpolation has a closed form: with a zero advector and
euler_a forcing R = a + b t (solvers/mpdata_rhs.hpp:135-183) the state is
psi0 + a t + b (t^2 - t dt) / 2 exactly at every step, a quadratic in time.
The quadratic order reproduces it at every record time, the linear order
through steps t0 < T <= t1 gives it plus -(b/2)(T - t0)(T - t1) (zero where
a record time is a step time), and with b = 0 both reproduce it.  The
schedule follows the reference's counters: a record per multiple of outfreq
crossed, written after the crossing step (order 1) or one step later
(order 2)."""
import numpy as np
import pytest

import oracle as O

GRID = (12, 10)
DT, OUTFREQ, DURATION = 0.25, 0.625, 3.0


def _run(ord_, a, b):
    o = O.Oracle(GRID, opts=O.FCT, n_iters=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=DT, var_dt=True, max_courant=0.5)
    psi0 = 1.0 + np.random.default_rng(5).random(GRID)
    o.set_advectee(psi0)
    for d in range(2):
        o.set_advector(np.zeros(o.advector_shape(d)), d)

    def upd(dt, at):
        o.set_rhs(0, np.full(GRID, a + b * o.time))
    o.set_rhs_scheme(O.RHS_EULER_A, upd)
    recs = []

    def out(kind, t):
        recs.append((kind, t, o.timestep, o.time, o.output(0)))
    o.set_var_dt_output(out, OUTFREQ, ord_)
    o.advance_time(DURATION)
    if ord_ == 2:
        o.advance(1)      # out_intrp_ord - 1 further steps (solver_common.hpp:293-294,352)
    return psi0, recs, o


def _truth(psi0, a, b, t):
    return psi0 + a * t + b * (t * t - t * DT) / 2


@pytest.mark.parametrize("ord_", [1, 2])
def test_schedule(ord_):
    psi0, recs, o = _run(ord_, 0.5, 0.0)
    assert o.timestep == 12 + (ord_ - 1)
    kinds = [r[0] for r in recs]
    assert kinds == [0, 1, 1, 1, 1]
    # start: t = end time of the call, the buffer holds the initial state
    assert recs[0][1] == DURATION and recs[0][2] == 0 and np.array_equal(recs[0][4], psi0)
    assert [r[1] for r in recs[1:]] == [0.625, 1.25, 1.875, 2.5]
    # crossing steps 3, 5, 8, 10 (of 0.25 from 0); order 2 records one step later
    assert [r[2] for r in recs[1:]] == [s + ord_ - 1 for s in (3, 5, 8, 10)]


@pytest.mark.parametrize("ord_", [1, 2])
def test_affine_in_time_is_exact(ord_):
    a = 0.5
    psi0, recs, _ = _run(ord_, a, 0.0)
    for _, t, _, _, y in recs[1:]:
        assert np.max(np.abs(y - _truth(psi0, a, 0.0, t))) <= 1e-13


def test_quadratic_in_time():
    a, b = 0.5, 0.75
    psi0, r2, _ = _run(2, a, b)
    _, r1, _ = _run(1, a, b)
    for _, t, _, _, y in r2[1:]:
        assert np.max(np.abs(y - _truth(psi0, a, b, t))) <= 1e-13
    for (_, t, ts, t1, y) in r1[1:]:
        t0 = t1 - DT
        want = _truth(psi0, a, b, t) - b / 2 * (t - t0) * (t - t1)
        assert np.max(np.abs(y - want)) <= 1e-13
        # the linear order is off the quadratic unless T is a step time
        assert (np.max(np.abs(y - _truth(psi0, a, b, t))) > 1e-3) == (t not in (t0, t1))


def test_rejections():
    o = O.Oracle(GRID, opts=O.FCT, n_iters=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=DT)
    with pytest.raises(ValueError):
        o.set_var_dt_output(lambda k, t: None, OUTFREQ, 1)     # constant dt
    v = O.Oracle(GRID, opts=O.FCT, n_iters=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=DT, var_dt=True)
    for bad in [dict(outfreq=0.0), dict(intrp_ord=3), dict(outvars=(1,))]:
        kw = dict(outfreq=OUTFREQ, intrp_ord=1, outvars=(0,))
        kw.update(bad)
        with pytest.raises(ValueError):
            v.set_var_dt_output(lambda k, t: None, **kw)

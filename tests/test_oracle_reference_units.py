"""The oracle on the known answers of the reference's own FCT unit tests
(tests/unit/cone_bugs/test_issue.cpp, tests/unit/hint_scale/hint_scale.cpp;
inputs in tests/reference_units.py)."""
import numpy as np
import pytest

import oracle as O
from reference_units import F32_EPS, cone_bugs_case, hint_scale_case


def run(case, opts, hint=0):
    o = O.Oracle(case["grid"], opts=opts, n_iters=case["n_iters"], bcs=case["bcs"],
                 hint_scale=[hint], max_abs_div_eps=F32_EPS)
    o.set_advectee(case["psi"])
    for d in range(len(case["grid"])):
        o.set_advector(case["gc"][d], d)
    o.advance(case["nt"])
    return o.advectee()


def test_cone_bugs_minimum_stays_one():
    # test_issue.cpp:100-101: slv.min() - 1 == 0 after 2 steps
    c = cone_bugs_case()
    assert c["psi"].min() == 1.0 and c["psi"].max() > 4.0
    out = run(c, O.FCT)
    assert out.min() == 1.0
    # control: the same cone without the limiter dips below the background
    assert run(c, 0).min() < 1.0


@pytest.mark.parametrize("hint", [0, 24])
def test_hint_scale_keeps_bounds(hint):
    # hint_scale.cpp:56-59: max <= initial max and min >= 0 after one step
    c = hint_scale_case()
    out = run(c, O.FCT | O.IGA, hint)
    assert np.isfinite(out).all()
    assert out.max() <= c["psi"].max()
    assert out.min() >= 0.0


def test_basic_example_log(golden_dir):
    # paper_2015_GMD/0_basic_example: the six printed records (5 decimals)
    from reference_units import basic_example_case, basic_example_records
    c = basic_example_case()
    want = basic_example_records(golden_dir)
    o = O.Oracle(c["grid"], opts=O.IGA | O.FCT, n_iters=c["n_iters"], bcs=c["bcs"])
    o.set_advectee(c["psi"])
    for d in range(2):
        o.set_advector(c["gc"][d], d)
    for k, rec in enumerate(want):
        if k:
            o.advance(c["outfreq"])
        got = o.advectee()[:, 1]
        assert np.abs(got - rec).max() <= 5e-6 + 1e-12, k
    # control: the basic scheme (no iga / fct) prints other numbers
    b = O.Oracle(c["grid"], opts=0, n_iters=2, bcs=c["bcs"])
    b.set_advectee(c["psi"])
    for d in range(2):
        b.set_advector(c["gc"][d], d)
    b.advance(c["nt"])
    assert np.abs(b.advectee()[:, 1] - want[-1]).max() > 1e-4


@pytest.mark.parametrize("group,opts", [(0, O.ABS), (1, O.IGA), (3, O.IGA | O.FCT)])
def test_advscheme_opts_log(golden_dir, group, opts):
    # paper_2015_GMD/1_advscheme_opts: both equations after 1200 steps, as
    # printed (5 decimals), for abs, iga and iga|fct
    from reference_units import advscheme_opts_case, advscheme_opts_finals
    c = advscheme_opts_case()
    want = advscheme_opts_finals(golden_dir, group)
    o = O.Oracle(c["grid"], opts=opts, n_iters=c["n_iters"], n_eqns=2, bcs=c["bcs"])
    for e in range(2):
        o.set_advectee(c["psi"][e], e)
    for d in range(2):
        o.set_advector(c["gc"][d], d)
    o.advance(c["nt"])
    for e in range(2):
        assert np.abs(o.advectee(e)[:, 1] - want[e]).max() <= 5e-6 + 1e-12, e

"""The oracle's explicit forcing (solvers::mpdata_rhs, solvers/mpdata_rhs.hpp)
pinned to the reference's coupled harmonic oscillator test
(tests/paper_2015_GMD/6_coupled_harmosc, trapezoidal scheme, refdata/
stats.txt: min/max with their indices to 8 decimals and the rms error every
100 steps), plus closed forms for the Euler schemes."""
import os

import numpy as np
import pytest

import oracle as O
from libmpdataxx_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "coupled_harmosc", "stats.txt")


def harmosc_oracle():
    psi, phi, gc, omega = W.coupled_harmosc()
    o = O.Oracle(psi.shape, opts=O.IGA | O.FCT, n_iters=2, n_eqns=2, bcs=[(O.CYCLIC, O.CYCLIC)] * 2, dt=1.0)
    o.set_advectee(psi, 0)
    o.set_advectee(phi, 1)
    for d in range(2):
        o.set_advector(gc[d], d)

    def upd(dt, at):
        r0, r1 = W.coupled_harmosc_rhs(o.advectee(0), o.advectee(1), dt, at, omega)
        o.set_rhs(0, r0)
        o.set_rhs(1, r1)
    o.set_rhs_scheme(O.RHS_TRAPEZ, upd)
    return o


def test_coupled_harmosc_matches_refdata():
    want = W.parse_coupled_harmosc_stats(open(GOLDEN).read())
    assert len(want) == 14
    o = harmosc_oracle()
    for ref in want:
        o.advance(ref["timestep"] - o.timestep)
        psi, phi = o.advectee(0), o.advectee(1)
        assert np.array_equal(psi[:, 0:1].repeat(psi.shape[1], 1), psi)   # stays y-uniform
        got = W.coupled_harmosc_stats(psi[:, 0], phi[:, 0], ref["timestep"])
        for k in ("min_psi", "max_psi", "min_phi", "max_phi"):
            assert f"{got[k][0]:.8f}".replace("-0.00000000", "0.00000000") == \
                f"{ref[k][0]:.8f}".replace("-0.00000000", "0.00000000"), (ref["timestep"], k)
            if abs(ref[k][0]) > 1e-6:
                assert got[k][1] == ref[k][1], (ref["timestep"], k)
        assert abs(got["rms"] - ref["rms"]) <= 5e-9 * abs(ref["rms"]), (ref["timestep"], got["rms"], ref["rms"])


@pytest.mark.parametrize("scheme", [O.RHS_EULER_A, O.RHS_EULER_B])
def test_euler_schemes_closed_form(scheme):
    """With a zero advector the step is the identity on advection, so a
    constant forcing R makes psi^{n+1} = psi^n + dt R for both Euler schemes
    (euler_a adds it after, euler_b before the advection), and a linear decay
    R = -k psi gives the explicit-Euler factor (1 - k dt)^n."""
    grid = (10, 9)
    psi0 = W.random_field(grid, 3)
    o = O.Oracle(grid, opts=O.FCT, n_iters=2, n_eqns=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=0.5)
    o.set_advectee(psi0, 0)
    o.set_advectee(psi0, 1)
    for d in range(2):
        o.set_advector(np.zeros(o.advector_shape(d)), d)
    k = 0.3

    def upd(dt, at):
        assert dt == 0.5 and at == 0
        o.set_rhs(0, np.full(grid, 2.0))
        o.set_rhs(1, -k * o.advectee(1))
    o.set_rhs_scheme(scheme, upd)
    o.advance(6)
    assert np.allclose(o.advectee(0), psi0 + 6 * 0.5 * 2.0, rtol=0, atol=1e-13)
    assert np.allclose(o.advectee(1), psi0 * (1 - k * 0.5) ** 6, rtol=1e-14, atol=0)

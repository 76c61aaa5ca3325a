"""GPU explicit forcing (solvers::mpdata_rhs schemes) against the reference's
coupled harmonic oscillator refdata and against the oracle."""
import os

import numpy as np
import pytest

import oracle as O
from gpu_util import rel_err, setup_case
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "coupled_harmosc", "stats.txt")


@pytest.mark.parametrize("path", [1, 2])
def test_coupled_harmosc_matches_refdata(path):
    import libmpdataxx_amd as M
    print(f"coupled harmonic oscillator path={path}", flush=True)
    psi, phi, gc, omega = W.coupled_harmosc()
    s = M.Solver(psi.shape, opts=M.IGA | M.FCT, n_iters=2, n_eqns=2, bcs=[(M.CYCLIC, M.CYCLIC)] * 2,
                 dt=1.0, path=path)
    s.set_advectee(psi, 0)
    s.set_advectee(phi, 1)
    for d in range(2):
        s.set_advector(gc[d], d)

    def upd(dt, at):
        r0, r1 = W.coupled_harmosc_rhs(s.advectee(0), s.advectee(1), dt, at, omega)
        s.set_rhs(0, r0)
        s.set_rhs(1, r1)
    s.set_rhs_scheme(M.RHS_TRAPEZ, upd)
    want = W.parse_coupled_harmosc_stats(open(GOLDEN).read())
    for ref in want:
        s.advance(ref["timestep"] - s.timestep)
        got = W.coupled_harmosc_stats(s.advectee(0)[:, 0], s.advectee(1)[:, 0], ref["timestep"])
        for k in ("min_psi", "max_psi", "min_phi", "max_phi"):
            assert f"{got[k][0]:.8f}".replace("-0.00000000", "0.00000000") == \
                f"{ref[k][0]:.8f}".replace("-0.00000000", "0.00000000"), (ref["timestep"], k)
        assert abs(got["rms"] - ref["rms"]) <= 5e-9 * abs(ref["rms"])
    s.close()


@pytest.mark.parametrize("scheme", [O.RHS_EULER_A, O.RHS_EULER_B, O.RHS_TRAPEZ])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rhs_schemes_match_oracle(dtype, scheme):
    """A nonlinear forcing R = 0.05 - 0.2 psi^2 (R^{n+1} evaluated on the
    advected state for trapez) on a 3D grid with interior tiles, two
    equations with different forcing signs."""
    import libmpdataxx_amd as M
    grid, bcs = (30, 26, 40), [(1, 1), (0, 0), (1, 1)]
    psi, gc = setup_case(grid, bcs, 7)
    print(f"rhs scheme={scheme} dtype={np.dtype(dtype).name}", flush=True)
    o = O.Oracle(grid, opts=O.FCT, n_iters=3, n_eqns=2, bcs=bcs, dt=0.5, max_abs_div_eps=1e300)
    s = M.Solver(grid, opts=M.FCT, n_iters=3, n_eqns=2, bcs=bcs, dt=0.5, dtype=dtype, max_abs_div_eps=1e300)
    for x in (o, s):
        x.set_advectee(psi, 0)
        x.set_advectee(psi[::-1].copy(), 1)
        for d in range(3):
            x.set_advector(gc[d], d)

    def forcing(x):
        def upd(dt, at):
            for e, sign in ((0, 1.0), (1, -1.0)):
                a = x.advectee(e).astype(np.float64)
                x.set_rhs(e, sign * (0.05 - 0.2 * a * a))
        return upd
    o.set_rhs_scheme(scheme, forcing(o))
    s.set_rhs_scheme(scheme, forcing(s))
    o.advance(4)
    s.advance(4)
    for e in range(2):
        assert rel_err(s.advectee(e).astype(np.float64), o.advectee(e)) <= (1e-12 if dtype == np.float64 else 1e-5)
    s.close()


@pytest.mark.parametrize("scheme", [O.RHS_EULER_A, O.RHS_EULER_B, O.RHS_TRAPEZ])
def test_mixed_scheme_hooks_reproduce_each_scheme(scheme):
    """The mixed scheme (rhs_scheme_t::mixed): its three hooks, written to do
    what mpdata_rhs.hpp does for euler_a / euler_b / trapez with the same
    update_rhs / apply_rhs calls, give bit for bit the fields of that scheme."""
    import libmpdataxx_amd as M
    grid, bcs = (24, 20, 30), [(0, 0), (1, 1), (0, 0)]
    psi, gc = setup_case(grid, bcs, 17)
    print(f"mixed scheme as scheme={scheme}", flush=True)

    def make():
        s = M.Solver(grid, opts=M.FCT, n_iters=2, n_eqns=1, bcs=bcs, dt=0.5, max_abs_div_eps=1e300)
        s.set_advectee(psi)
        for d in range(3):
            s.set_advector(gc[d], d)
        return s

    def forcing(s):
        def upd(dt, at):
            a = s.advectee(0)
            s.set_rhs(0, 0.05 - 0.2 * a * a + 0.01 * at)
        return upd

    ref = make()
    ref.set_rhs_scheme(scheme, forcing(ref))
    ref.advance(4)
    want = ref.advectee()
    ref.close()

    s = make()
    s.set_rhs_scheme(M.RHS_MIXED, forcing(s))
    dt = 0.5
    calls = []

    def hook(which):
        calls.append(which)
        if which == M.HOOK_ANTE_LOOP and scheme == O.RHS_TRAPEZ:
            s.update_rhs(dt / 2, 0)
        elif which == M.HOOK_ANTE_STEP:
            if scheme in (O.RHS_EULER_A, O.RHS_EULER_B):
                s.update_rhs(dt, 0)
            if scheme == O.RHS_EULER_B:
                s.apply_rhs(dt)
            if scheme == O.RHS_TRAPEZ:
                s.apply_rhs(dt / 2)
        elif which == M.HOOK_POST_STEP:
            if scheme == O.RHS_EULER_A:
                s.apply_rhs(dt)
            if scheme == O.RHS_TRAPEZ:
                s.update_rhs(dt / 2, 1)
                s.apply_rhs(dt / 2)
    s.set_rhs_hooks(hook)
    s.advance(4)
    got = s.advectee()
    s.close()
    assert calls == [M.HOOK_ANTE_LOOP] + [M.HOOK_ANTE_STEP, M.HOOK_POST_STEP] * 4
    assert np.array_equal(got, want)

"""GPU runs of the reference's own regression problems.

3D revolving sphere (tests/paper_2015_GMD/4_revolving_sphere_3d, 59^3, 556
steps, FP64): the GPU fields must match the reference's float32 refdata and
its 8-decimal statistics, like the oracle does.  Also checks the error path
of the divergence check (solver_3d.hpp:375-388) and the panic flag.
"""
import gzip
import os

import numpy as np
import pytest

import libmpdataxx_amd as M
from libmpdataxx_amd import workloads as W

pytestmark = pytest.mark.gpu

SPHERE = {"upwind": (0, 1), "basic": (0, 2), "iga": (M.IGA, 2),
          "iga_fct": (M.IGA | M.FCT, 2), "fct": (M.FCT, 2)}


def read_stats(path):
    lines = [l for l in open(path).read().splitlines() if "=" in l]
    return {k.strip(): float(v) for k, v in (l.split("=") for l in lines[3:])}


@pytest.mark.parametrize("variant", list(SPHERE))
def test_revolving_sphere_gpu_fp64(golden_dir, variant):
    print(f"revolving sphere {variant}", flush=True)
    opts, iters = SPHERE[variant]
    psi, gc, dt = W.revolving_sphere_3d()
    s = M.Solver((59, 59, 59), opts=opts, n_iters=iters, dt=dt)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    s.advance(556)
    fc = s.advectee()
    assert s.timestep == 556 and abs(s.time - 556 * dt) < 1e-9
    ref = np.frombuffer(gzip.open(os.path.join(golden_dir, "revolving_sphere_3d", f"{variant}_556.f32.gz")).read(),
                        dtype="<f4").reshape(59, 59, 59).astype(np.float64)
    tol = np.maximum(np.abs(ref) * 2.0 ** -23, 1e-12)
    assert np.all(np.abs(fc - ref) <= tol)
    st = W.revolving_sphere_stats(psi, fc, 556, dt)
    want = read_stats(os.path.join(golden_dir, "revolving_sphere_3d", f"stats_{variant}.txt"))
    for k, w in (("max", "max(solution)"), ("min", "min(solution)"), ("Linf", "Linf"), ("L2", "L2"), ("L1", "L1")):
        assert f"{st[k]:.8f}" == f"{want[w]:.8f}", k
    assert s.min() == fc.min() and s.max() == fc.max()


def test_divergent_advector_throws():
    s = M.Solver((8, 8, 8), opts=M.FCT, n_iters=2)
    s.set_advectee(np.ones((8, 8, 8)))
    gx = np.zeros((8, 8, 8))
    gx[3, 4, 4] = 0.1
    s.set_advector(gx, 0)
    s.set_advector(np.zeros((9, 7, 8)), 1)
    s.set_advector(np.zeros((9, 8, 7)), 2)
    with pytest.raises(M.MpdxError, match="initial advector field is divergent"):
        s.advance(1)


def test_panic_flag_stops_advance():
    psi, gc, dt = W.revolving_sphere_3d(21)
    s = M.Solver((21, 21, 21), opts=M.FCT, n_iters=2, dt=dt)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    s.advance(2)
    s.set_panic(True)
    s.advance(5)
    assert s.timestep == 2
    s.set_panic(False)
    s.advance(3)
    assert s.timestep == 5

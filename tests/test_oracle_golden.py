"""Pins the FP64 CPU oracle against the reference's own golden data.

* 3D revolving sphere (tests/paper_2015_GMD/4_revolving_sphere_3d): the
  reference's float32 fields after 556 steps for upwind/basic/iga/iga_fct/fct
  and its 8-decimal statistics (revolving_sphere_stats.hpp:46-65).
* 2D rotating cone (tests/paper_2015_GMD/3_rotating_cone_2d): the
  reference's min/max/rms statistics after 3768 steps.
"""
import gzip
import os

import numpy as np
import pytest

import oracle as O
from libmpdataxx_amd import workloads as W

SPHERE = {"upwind": (0, 1), "basic": (0, 2), "iga": (O.IGA, 2),
          "iga_fct": (O.IGA | O.FCT, 2), "fct": (O.FCT, 2)}
CONE = {"basic": (0, 2), "fct": (O.FCT, 2), "iga_fct": (O.IGA | O.FCT, 2)}


def load_f32(path):
    return np.frombuffer(gzip.open(path).read(), dtype="<f4").reshape(59, 59, 59)


def read_stats(path):
    out = {}
    lines = [l for l in open(path).read().splitlines() if "=" in l]
    for l in lines[3:]:          # skip the t=0 block
        k, v = (x.strip() for x in l.split("="))
        out[k] = float(v)
    return out


def test_sphere_initial_condition(golden_dir):
    psi, _, _ = W.revolving_sphere_3d()
    ref = load_f32(os.path.join(golden_dir, "revolving_sphere_3d", "initial_0.f32.gz"))
    assert np.array_equal(psi.astype(np.float32), ref)


@pytest.mark.parametrize("variant", list(SPHERE))
def test_revolving_sphere_fields_and_stats(golden_dir, variant):
    opts, iters = SPHERE[variant]
    psi, gc, dt = W.revolving_sphere_3d()
    o = O.Oracle((59, 59, 59), opts=opts, n_iters=iters, dt=dt)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(556)
    fc = o.advectee()
    ref = load_f32(os.path.join(golden_dir, "revolving_sphere_3d", f"{variant}_556.f32.gz"))
    # the reference wrote float32 of its FP64 field: agreement to well below
    # one float32 ulp of the field scale (4.0) means the FP64 trajectories agree
    diff = np.abs(fc - ref.astype(np.float64))
    tol = np.maximum(np.abs(ref).astype(np.float64) * 2.0 ** -23, 1e-12)
    assert np.all(diff <= tol), (diff.max(), int((diff > tol).sum()))
    stats = W.revolving_sphere_stats(psi, fc, 556, dt)
    want = read_stats(os.path.join(golden_dir, "revolving_sphere_3d", f"stats_{variant}.txt"))
    assert f"{stats['max']:.8f}" == f"{want['max(solution)']:.8f}"
    assert f"{stats['min']:.8f}" == f"{want['min(solution)']:.8f}"
    assert f"{stats['Linf']:.8f}" == f"{want['Linf']:.8f}"
    assert f"{stats['L2']:.8f}" == f"{want['L2']:.8f}"
    assert f"{stats['L1']:.8f}" == f"{want['L1']:.8f}"


@pytest.mark.parametrize("variant", list(CONE))
def test_rotating_cone_stats(golden_dir, variant):
    opts, iters = CONE[variant]
    psi, gc, dt = W.rotating_cone_2d()
    o = O.Oracle((101, 101), opts=opts, n_iters=iters, dt=dt)
    o.set_advectee(psi)
    for d in range(2):
        o.set_advector(gc[d], d)
    o.advance(628 * 6)
    stats = W.rotating_cone_stats(psi, o.advectee(), 3768, dt)
    want = read_stats(os.path.join(golden_dir, "rotating_cone_2d", f"stats_{variant}_.txt"))
    assert f"{stats['min']:.8f}" == f"{want['min(psi)']:.8f}"
    assert f"{stats['max']:.8f}" == f"{want['max(psi)']:.8f}"
    assert f"{stats['rms']:.8f}" == f"{want['rms error']:.8f}"

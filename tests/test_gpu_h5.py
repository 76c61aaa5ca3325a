"""HDF5 / XDMF output of GPU runs: the reference's revolving-sphere program
with output::hdf5_xdmf, rewritten against the C++ API, writes the refdata
layout (compared with the reference's files), and the Python decorator
records the states a plain run passes through at the scheduled steps."""
import gzip
import hashlib
import os
import subprocess

import numpy as np
import pytest

import libmpdataxx_amd as M
from libmpdataxx_amd import build as B
from libmpdataxx_amd import h5 as H

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not B.have_hdf5(), reason="HDF5 is not installed")]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = os.path.join(B.HDF5_ROOT, "bin", "h5dump")
NX = 59
VARIANTS = ["upwind", "basic", "iga", "iga_fct", "fct"]


def golden_field(golden_dir, name):
    raw = gzip.open(os.path.join(golden_dir, "revolving_sphere_3d", name)).read()
    return np.frombuffer(raw, dtype="<f4").reshape(NX, NX, NX)


def test_revolving_sphere_cpp_output(tmp_path, golden_dir):
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "revolving_sphere_3d_h5")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m libmpdataxx_amd.build)")
    print("running revolving_sphere_3d_h5", flush=True)
    subprocess.run([exe, str(tmp_path)], check=True, timeout=300)
    g = os.path.join(golden_dir, "revolving_sphere_3d_h5")
    sums = {}
    for line in open(os.path.join(g, "const_data.sha256")):
        digest, key = line.split()
        sums[key] = digest
    for v in VARIANTS:
        out = tmp_path / v
        for name in ("timestep0000000000.xmf", "timestep0000000556.xmf", "temp.xmf"):
            assert open(out / name, "rb").read() == open(os.path.join(g, name), "rb").read(), (v, name)
        dump = subprocess.run([H5DUMP, "-A", "const.h5"], cwd=out, capture_output=True, check=True).stdout
        assert dump == open(os.path.join(g, f"const_{v}.txt"), "rb").read(), v
        for ds, shape in (("X", (NX + 1,) * 3), ("Y", (NX + 1,) * 3), ("Z", (NX + 1,) * 3), ("T", (2,))):
            a = H.read(str(out / "const.h5"), ds, shape)
            assert hashlib.sha256(a.astype("<f4").tobytes()).hexdigest() == sums[f"{v}/const.h5:/{ds}"], (v, ds)
        # the initial condition is computed on the host: bit-identical
        np.testing.assert_array_equal(H.read(str(out / "timestep0000000000.h5"), "psi", (NX,) * 3),
                                      golden_field(golden_dir, "initial_0.f32.gz"))
        # the FP64 GPU result stored as float32: within 2^-22 of the
        # reference's own stored result (as the field comparison of the
        # stats programs)
        f = H.read(str(out / "timestep0000000556.h5"), "psi", (NX,) * 3).astype(np.float64)
        ref = golden_field(golden_dir, f"{v}_556.f32.gz").astype(np.float64)
        assert np.all(np.abs(f - ref) <= np.maximum(np.abs(ref) * 2.0 ** -22, 1e-12)), v


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_python_decorator_records_scheduled_states(tmp_path, dtype):
    from libmpdataxx_amd import workloads as W

    grid = (24, 20, 22)
    psi = W.random_field(grid, 11)
    chi = W.random_field(grid, 12)
    gc = W.streamfunction_advector(grid, 13, periodic=False, max_courant=0.4)

    def make():
        s = M.Solver(grid, opts=M.FCT, n_iters=3, n_eqns=2, dtype=dtype, dt=0.5)
        s.set_advectee(psi, 0)
        s.set_advectee(chi, 1)
        for d in range(3):
            s.set_advector(gc[d], d)
        return s

    nt, outfreq, outwindow = 7, 3, 2
    want = {}
    s = make()
    want[0] = (s.advectee(0), s.advectee(1))
    for ts in range(1, nt + 1):
        s.advance(1)
        want[ts] = (s.advectee(0), s.advectee(1))
    s.close()

    s = make()
    out = H.Hdf5Xdmf(s, tmp_path, outfreq=outfreq, outwindow=outwindow, outvars={0: "psi", 1: "chi"},
                     dijk=(0.5, 0.5, 0.5))
    out.advance(nt)
    assert out.records == [0, 1, 3, 4, 6, 7]
    t = H.read(str(tmp_path / "const.h5"), "T", (6,))
    np.testing.assert_array_equal(t, np.array([0, 1, 3, 4, 6, 7], dtype=np.float32) * np.float32(0.5))
    for ts in out.records:
        path = str(tmp_path / f"timestep{ts:010d}.h5")
        np.testing.assert_array_equal(H.read(path, "psi", grid), want[ts][0].astype(np.float32))
        np.testing.assert_array_equal(H.read(path, "chi", grid), want[ts][1].astype(np.float32))
    s.close()


def test_python_decorator_two_advance_calls(tmp_path):
    """advance(4) then advance(3) records what advance(7) records: the start
    record and const.h5 come only from the call at timestep 0 (hook_ante_loop,
    which solver_common.hpp:277-285 runs only then), so const.h5's time axis
    is the first call's (the records of an advance over 4 steps)."""
    from libmpdataxx_amd import workloads as W

    grid = (24, 20, 22)
    psi = W.random_field(grid, 21)
    gc = W.streamfunction_advector(grid, 22, periodic=False, max_courant=0.4)
    s = M.Solver(grid, opts=M.FCT, n_iters=2, dtype=np.float64, dt=0.5)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    out = H.Hdf5Xdmf(s, tmp_path, outfreq=3, outwindow=2, dijk=(0.5, 0.5, 0.5))
    out.advance(4)
    mid = s.advectee()
    out.advance(3)
    assert out.records == [0, 1, 3, 4, 6, 7]
    assert s.timestep == 7
    t = H.read(str(tmp_path / "const.h5"), "T", (4,))
    np.testing.assert_array_equal(t, np.array([0, 1, 3, 4], dtype=np.float32) * np.float32(0.5))
    np.testing.assert_array_equal(H.read(str(tmp_path / "timestep0000000004.h5"), "psi", grid),
                                  mid.astype(np.float32))
    s.close()

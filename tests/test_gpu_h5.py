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


def _oracle_var_dt_records(init, grid, ord_, durations, outfreq, **kw):
    """The oracle's var_dt output (start and records: kind, time, timestep,
    interpolated field) for the given initial advectee and advector."""
    import oracle as O
    o = O.Oracle(grid, **kw)
    o.set_advectee(init[0])
    for d in range(len(grid)):
        o.set_advector(init[1 + d], d)
    recs = []
    o.set_var_dt_output(lambda kind, t: recs.append((kind, t, o.timestep, o.output(0))), outfreq, ord_)
    for dur in durations:
        o.advance_time(dur)
        if ord_ == 2:
            o.advance(1)
    return recs, o


def _check_var_dt_files(outdir, recs, grid, name="psi"):
    import re
    assert [r[0] for r in recs] == [0] + [1] * (len(recs) - 1) and len(recs) >= 4
    temp = open(os.path.join(outdir, "temp.xmf")).read()
    listed = re.findall(r"timestep(\d{10})\.xmf", temp)
    assert [int(x) for x in listed] == [r[2] for r in recs]
    for kind, t, ts, want in recs:
        base = os.path.join(outdir, f"timestep{ts:010d}")
        xmf = open(base + ".xmf").read()
        # the start record is at the initial time (its callback's t is the
        # first call's end time)
        rt = t if kind == 1 else 0.0
        assert float(re.search(r'<Time Value="([^"]+)"', xmf).group(1)) == pytest.approx(rt, abs=1e-6)
        got = H.read(base + ".h5", name, grid).astype(np.float64)
        assert np.all(np.abs(got - want) <= np.abs(want) * 2.0 ** -22 + 1e-12), (ts, t)


def test_var_dt_cpp_output(tmp_path):
    """output::hdf5_xdmf over a var_dt solver through the C++ API
    (tests/cpp/var_dt_h5.cpp, out_intrp_ord 1 and 2): the start record and the
    records at k * outfreq, time-interpolated (output_common.hpp:47-62,
    72-140), hold the oracle's interpolated fields (float32 storage, 2^-22),
    are named by the timestep they were written after and carry their record
    time in the XDMF markup; const.h5's time axis covers the first call
    (var_dt: multiples of outfreq up to its end time, hdf5.hpp:199-210)."""
    import oracle as O
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "var_dt_h5")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m libmpdataxx_amd.build)")
    print("running var_dt_h5", flush=True)
    subprocess.run([exe, str(tmp_path)], check=True, timeout=300)
    grid = (40, 30)
    probe = O.Oracle(grid, opts=O.FCT, n_iters=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=0.3)
    shapes = [grid] + [probe.advector_shape(d) for d in range(2)]
    for ord_ in (1, 2):
        out = tmp_path / f"ord{ord_}"
        raw = np.fromfile(out / "in.f64", dtype=np.float64)
        parts = np.split(raw, np.cumsum([int(np.prod(s)) for s in shapes])[:-1])
        assert sum(p.size for p in parts) == raw.size
        init = [p.reshape(s) for p, s in zip(parts, shapes)]
        recs, o = _oracle_var_dt_records(init, grid, ord_, (1.3, 0.9), 0.5, opts=O.FCT, n_iters=2,
                                         bcs=[(O.OPEN, O.OPEN)] * 2, dt=0.3, var_dt=True, max_courant=0.4)
        _check_var_dt_files(str(out), recs, grid)
        np.testing.assert_array_equal(H.read(str(out / "const.h5"), "T", (3,)),
                                      np.array([0.0, 0.5, 1.0], dtype=np.float32))
        dump = subprocess.run([H5DUMP, "-A", "const.h5"], cwd=out, capture_output=True, check=True).stdout.decode()
        assert "max_courant" in dump and "var_dt" in dump


def test_python_decorator_var_dt(tmp_path):
    """The Python decorator over a var_dt solver (3D, two equations, FP64,
    out_intrp_ord 2): files hold the oracle's interpolated records."""
    import oracle as O
    from libmpdataxx_amd import workloads as W

    grid = (24, 20, 22)
    psi = W.random_field(grid, 31)
    gc = W.streamfunction_advector(grid, 32, periodic=False, max_courant=0.4)
    kw = dict(opts=M.FCT, n_iters=3, dt=0.5, var_dt=True, max_courant=0.45)
    s = M.Solver(grid, dtype=np.float64, **kw)
    s.set_advectee(psi)
    for d in range(3):
        s.set_advector(gc[d], d)
    out = H.Hdf5Xdmf(s, tmp_path, outfreq=0.75, intrp_ord=2, dijk=(0.5, 0.5, 0.5))
    out.advance(2.0)
    out.advance(1.1)
    recs, o = _oracle_var_dt_records([psi] + list(gc), grid, 2, (2.0, 1.1), 0.75, bcs=[(O.OPEN, O.OPEN)] * 3,
                                     **kw)
    assert s.timestep == o.timestep
    assert out.records == [r[2] for r in recs]
    _check_var_dt_files(str(tmp_path), recs, grid)
    s.close()

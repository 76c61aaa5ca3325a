"""The reference's own regression programs, rewritten against the C++ API
(include/libmpdata++/concurr/hip.hpp) and run on the GPU: their statistics
files must equal the reference's refdata statistics digit for digit (the
reference compares them with zdiff, tests/paper_2015_GMD/*/CMakeLists.txt),
and the 3D fields must match the refdata float32 fields."""
import gzip
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _run(name, outdir):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m libmpdataxx_amd.build)")
    print(f"running {name}", flush=True)
    subprocess.run([exe, str(outdir)], check=True, timeout=300)


def test_revolving_sphere_cpp(tmp_path, golden_dir):
    _run("revolving_sphere_3d", tmp_path)
    for v in ["upwind", "basic", "iga", "iga_fct", "fct"]:
        got = open(tmp_path / f"stats_{v}.txt").read().split()
        want = open(os.path.join(golden_dir, "revolving_sphere_3d", f"stats_{v}.txt")).read().split()
        assert got == want, v
        f = np.fromfile(tmp_path / f"{v}_556.f32", dtype=np.float32).astype(np.float64)
        ref = np.frombuffer(gzip.open(os.path.join(golden_dir, "revolving_sphere_3d", f"{v}_556.f32.gz")).read(),
                            dtype="<f4").astype(np.float64)
        assert np.all(np.abs(f - ref) <= np.maximum(np.abs(ref) * 2.0 ** -22, 1e-12)), v


def test_rotating_cone_cpp(tmp_path, golden_dir):
    _run("rotating_cone_2d", tmp_path)
    for v in ["basic", "fct", "iga_fct"]:
        got = open(tmp_path / f"stats_{v}_.txt").read().split()
        want = open(os.path.join(golden_dir, "rotating_cone_2d", f"stats_{v}_.txt")).read().split()
        assert got == want, v


@pytest.mark.parametrize("ord_", [1, 2])
def test_var_dt_cpp_out_intrp_ord(tmp_path, ord_):
    """A var_dt solver of the C++ API (tests/cpp/var_dt_steps.cpp) advanced by
    0, 1.3 and 0.25 with out_intrp_ord 1 and 2: each call steps while time <
    end or additional_steps > 0 (solver_common.hpp:292-355), so it takes at
    least one step and, with out_intrp_ord 2, one more once time has reached
    the end.  The oracle is driven the same way (advance by a duration, then
    one step for ord 2) from the initial state the program wrote; times agree
    to 1e-15 and the FP64 fields to 1e-12 relative."""
    import oracle as O
    _run("var_dt_steps", tmp_path)
    grid = (40, 30)
    o = O.Oracle(grid, opts=O.FCT, n_iters=2, bcs=[(O.OPEN, O.OPEN)] * 2, dt=0.3,
                 var_dt=True, max_courant=0.4)
    raw = np.fromfile(tmp_path / f"var_dt_ord{ord_}_in.f64", dtype=np.float64)
    sizes = [40 * 30] + [int(np.prod(o.advector_shape(d))) for d in range(2)]
    assert raw.size == sum(sizes)
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    o.set_advectee(parts[0].reshape(grid))
    for d in range(2):
        o.set_advector(parts[1 + d].reshape(o.advector_shape(d)), d)
    times = []
    for dur in (0.0, 1.3, 0.25):
        t0 = o.time
        o.advance_time(dur)
        if ord_ == 2:
            o.advance(1)
        assert o.time >= t0 + dur
        times.append(o.time)
    got_t = [float(x) for x in open(tmp_path / f"var_dt_ord{ord_}.txt").read().split()]
    assert np.allclose(got_t, times, rtol=1e-15, atol=0), (got_t, times)
    got = np.fromfile(tmp_path / f"var_dt_ord{ord_}.f64", dtype=np.float64).reshape(grid)
    want = o.advectee()
    assert np.sqrt(np.sum((got - want) ** 2) / np.sum(want ** 2)) <= 1e-12


def test_headline_through_cpp_api_at_device_speed():
    """The 512^3 FP32 MPDATA-FCT workload written as a libmpdata++ program
    with concurr::hip (tests/cpp/sphere_c3_bench.cpp): advance(K) is one
    library call with no host transfer once the inputs are on the GPU, so its
    step rate is within 5 % of the Python binding's device-buffer run of the
    same workload, and both give the same fields (the same FP64 mass after
    the same steps)."""
    import json
    import time

    import torch

    import libmpdataxx_amd as M
    from libmpdataxx_amd.workloads_torch import rotation_sphere_device
    n, K, W = 512, 20, 3
    exe = os.path.join(BIN, "sphere_c3_bench")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m libmpdataxx_amd.build)")
    print(f"C++ API headline {n}^3 K={K}", flush=True)
    r = subprocess.run([exe, str(n), str(K), str(W)], check=True, timeout=300, capture_output=True, text=True)
    cpp = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"  C++ {cpp['gcell_per_s']:.2f} Gcell/s", flush=True)

    dev = torch.device("cuda:0")
    s = M.Solver((n,) * 3, opts=M.FCT, n_iters=3, dtype=np.float32)
    psi, gc = rotation_sphere_device(n, torch.float32, dev)
    s.set_advectee_device(psi)
    for d in range(3):
        s.set_advector_device(gc[d], d)
    del psi, gc
    m0 = s.mass()
    s.advance(W)
    best = 1e300
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.advance(K)
        best = min(best, time.perf_counter() - t0)
    py = n ** 3 * K / best / 1e9
    print(f"  Python {py:.2f} Gcell/s", flush=True)
    assert cpp["mass0"] == m0
    assert cpp["mass"] == s.mass()
    s.close()
    assert cpp["gcell_per_s"] >= 0.95 * py

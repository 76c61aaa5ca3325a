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

"""Builds the HIP library libmpdx.so for gfx950 in-tree (no JIT cache).

    python -m libmpdataxx_amd.build

Packed FP32 instructions are disabled (-packed-fp32-ops): the compiler would
otherwise form v_pk_* FP32 ops from ordinary float code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpdx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPDX_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
SOURCES = ["mpdx_capi.hip"]


CPP_TESTS = ["revolving_sphere_3d", "rotating_cone_2d"]


def build_cpp_tests(verbose: bool = False) -> None:
    """Compile the C++ API regression programs of tests/cpp against libmpdx.so."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    for name in CPP_TESTS:
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
               os.path.join(src_dir, name + ".cpp"), "-L", HERE, "-lmpdx",
               "-Wl,-rpath,$ORIGIN/../../../libmpdataxx_amd", "-o", os.path.join(out_dir, name)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libmpdx.so; returns the library path."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_tests(verbose=True)
    sys.exit(0)

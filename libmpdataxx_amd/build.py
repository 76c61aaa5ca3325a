"""Builds the HIP library libmpdx.so for gfx950 in-tree (no JIT cache).

    python -m libmpdataxx_amd.build

Packed FP32 instructions are disabled (-packed-fp32-ops): the compiler would
otherwise form v_pk_* FP32 ops from ordinary float code.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpdx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPDX_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
# source -> extra flags: the FP32 interior kernels get the max-ILP
# instruction scheduler (their occupancy is fixed by LDS, not registers)
SOURCES = {
    "mpdx_capi.hip": [],
    "mpdx_interior.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


CPP_TESTS = ["revolving_sphere_3d", "rotating_cone_2d", "sphere_c3_bench", "var_dt_steps"]
# HDF5 (output only): the image's HDF5 C library; the writer is host code
HDF5_ROOT = os.environ.get("MPDX_HDF5_ROOT", "/opt/conda")
H5LIB = os.path.join(HERE, "libmpdx_h5.so")
CPP_TESTS_H5 = ["revolving_sphere_3d_h5", "var_dt_h5"]


def have_hdf5() -> bool:
    return os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h"))


def _h5_flags():
    # HDF5 by full path, not -L: that directory also holds an older libstdc++
    # which must not shadow the system one the ROCm runtime needs; for the
    # same reason the writer carries its own copy of libstdc++/libgcc.
    lib = os.path.join(HDF5_ROOT, "lib")
    return ["-I", os.path.join(HDF5_ROOT, "include"), os.path.join(lib, "libhdf5.so"),
            f"-Wl,-rpath,{lib}", "-static-libstdc++", "-static-libgcc", "-Wl,--exclude-libs,ALL"]


def build_h5(verbose: bool = False) -> str | None:
    """Compile the HDF5/XDMF writer libmpdx_h5.so (host code, g++) when the
    HDF5 headers are installed; returns its path or None."""
    if not have_hdf5():
        if verbose:
            print(f"HDF5 not found under {HDF5_ROOT}: libmpdx_h5.so not built", flush=True)
        return None
    tmp = H5LIB + ".tmp"
    with tempfile.TemporaryDirectory() as td:
        # export the C ABI only (not the template code of the bundled libstdc++)
        vs = os.path.join(td, "exports.map")
        with open(vs, "w") as f:
            f.write("{ global: mpdx_h5_*; local: *; };\n")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-fvisibility=hidden",
               os.path.join(CSRC, "mpdx_h5.cpp"), *_h5_flags(), f"-Wl,--version-script={vs}", "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, H5LIB)
    return H5LIB


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
    if have_hdf5():
        for name in CPP_TESTS_H5:
            cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                   os.path.join(src_dir, name + ".cpp"), "-L", HERE, "-lmpdx", "-lmpdx_h5",
                   "-Wl,-rpath,$ORIGIN/../../../libmpdataxx_amd", "-o", os.path.join(out_dir, name)]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)


RESOURCES = os.path.join(HERE, "kernel_resources.txt")


def _resource_report(stderr: str) -> str:
    """One line per kernel from the compiler's kernel-resource-usage remarks:
    name, VGPRs, AGPRs, scratch bytes per lane, VGPR and SGPR spills, LDS."""
    keys = {"VGPRs": "vgpr", "AGPRs": "agpr", "ScratchSize [bytes/lane]": "scratch",
            "VGPRs Spill": "vspill", "SGPRs Spill": "sspill", "LDS Size [bytes/block]": "lds"}
    rows, cur = [], None
    for line in stderr.splitlines():
        if "remark:" not in line or "kernel-resource-usage" not in line:
            continue
        body = line.split("remark:", 1)[1].split("[-Rpass", 1)[0].strip()
        if body.startswith("Function Name:"):
            cur = {"name": body.split(":", 1)[1].strip()}
            rows.append(cur)
        elif cur is not None and ":" in body:
            k, v = body.rsplit(":", 1)
            if k.strip() in keys:
                cur[keys[k.strip()]] = v.strip()
    out = []
    for r in rows:
        out.append(" ".join([r["name"]] + [f"{k}={r.get(k, '?')}" for k in keys.values()]))
    return "\n".join(out) + "\n"


def build(verbose: bool = False) -> str:
    """Compile every HIP source (each with its own flags, in parallel) and
    link libmpdx.so; returns the library path.  The compiler's per-kernel
    register / scratch / LDS usage is written to kernel_resources.txt next to
    the library."""
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory() as td:
        procs, objs = [], []
        for src, extra in SOURCES.items():
            obj = os.path.join(td, src + ".o")
            cmd = [HIPCC, *FLAGS, *extra, "-Rpass-analysis=kernel-resource-usage", "-c",
                   os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        report = []
        for cmd, p in procs:
            _, err = p.communicate()
            if p.returncode != 0:
                sys.stderr.write(err)
                raise subprocess.CalledProcessError(p.returncode, cmd)
            report.append(_resource_report(err))
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    with open(RESOURCES, "w") as f:
        f.write("".join(report))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    build_h5(verbose=True)
    build_cpp_tests(verbose=True)
    sys.exit(0)

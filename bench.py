#!/usr/bin/env python3
"""Benchmark of the MPDATA advection step on one MI355X.

Headline workload: 3D MPDATA with FCT, n_iters = 3, open boundaries, 512^3
grid, FP32 state/advector/fluxes with FP64 guards and FP64 Kahan mass sums;
solid-body rotation about z (max Courant 0.5) moving a discontinuous sphere.
A "step" is one time step of the whole grid.  The same problem is also run
in FP64 for the FP32/FP64 ratio and as the accuracy reference (the FP64 GPU
path agrees with the FP64 CPU oracle to <= 1e-12).  A second leg runs the
multi-tracer workload (512x512x128 cyclic, 8 tracers, time-varying
deformational advector generated on the device each step).  Inputs are
resident in HBM before the timed regions.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
METRIC = "Gcell-updates/s, 3D MPDATA-FCT iters=3, FP32 vs FP64, % of HBM peak"

# Compulsory HBM words per cell for one launch of each kernel class (reads
# of the arrays it consumes + writes of the arrays it produces; halo re-reads,
# values kept on chip and reductions excluded), and which cells one launch
# covers: "all" = n^3, "interior" = (n-4)^3 (cells >= 2 from every boundary).
ALG_WORDS = {
    "fused_interior_fct_v": (9, "interior"),  # psi^k, U x3, psi^n in; psi^k+1, V x3 out
    "fused_interior_fct": (6, "interior"),    # psi^k, U x3, psi^n in; psi^k+1 out
    "fused_interior_v": (8, "interior"),      # psi^k, U x3 in; psi^k+1, V x3 out
    "fused_interior": (5, "interior"),        # psi^k, U x3 in; psi^k+1 out
    "fused_first_fct_v": (8, "interior"),     # psi^n, GC x3 in; psi^2, V x3 out (psi^1 on chip)
    "fused_first_fct": (5, "interior"),       # psi^n, GC x3 in; psi^2 out
    "fused_first_v": (8, "interior"),
    "fused_first": (5, "interior"),
    "fused_donor": (5, "all"),                # psi^n, GC x3 in; psi^1 out
    "fused_psi_halo": (0, "all"),             # halo copies only
    "pp_antidiff": (5, "all"),                # psi, U_d, U_d+1, U_d+2 in; V_d out
    "pp_flux": (3, "all"),
    "pp_donorcell": (5, "all"),
    "pp_betas": (10, "all"),
    "pp_gc_mono": (5, "all"),
    "pp_fct_init": (3, "all"),
}


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def profile_figure(fname, key, kernel):
    """A per-launch figure recorded by rocprofv3 counter runs of this workload
    (machine-readable lines '<key> <timer name> <value>' in profiles/)."""
    try:
        with open(os.path.join(PROFILES, fname)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == key and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def run_gpu(torch, M, n, dtype, steps, warmup, path):
    dev = torch.device("cuda:0")
    s = M.Solver((n, n, n), opts=M.FCT, n_iters=3, dtype=dtype, path=path,
                 bcs=[(M.OPEN, M.OPEN)] * 3)
    from libmpdataxx_amd.workloads_torch import rotation_sphere_device
    psi, gc = rotation_sphere_device(n, torch.float32 if dtype == "float32" else torch.float64, dev)
    s.set_advectee_device(psi)
    for d in range(3):
        s.set_advector_device(gc[d], d)
    del psi, gc
    torch.cuda.synchronize()
    mass0 = s.mass()
    s.advance(warmup)
    torch.cuda.synchronize()
    # timed region: each step replays a captured hipGraph
    t0 = time.perf_counter()
    s.advance(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the same number of steps again, launched eagerly with a HIP event pair
    # around every kernel on the solver's stream (per-kernel durations)
    s.profile(True)
    t2 = time.perf_counter()
    s.advance(steps)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    prof = s.profile_read()
    s.profile(False)
    return s, (t1 - t0), prof, mass0, (t3 - t2)


def run_multi_tracer(torch, M, steps, warmup, shape=(512, 512, 128), n_tr=8):
    """3D MPDATA-FCT n_iters=3, cyclic, 8 tracers sharing one time-varying
    (reversing deformational) advector generated on the device every step."""
    dev = torch.device("cuda:0")
    s = M.Solver(shape, opts=M.FCT, n_iters=3, n_eqns=n_tr, dtype="float32", dt=1.0,
                 bcs=[(M.CYCLIC, M.CYCLIC)] * 3)
    from libmpdataxx_amd.workloads_torch import multi_tracer_fields
    for e, f in enumerate(multi_tracer_fields(shape, dev)):
        s.set_advectee_device(f, e)
    del f
    for d in range(3):
        ext = s.advector_shape(d)
        s.set_advector_device(torch.zeros(ext, dtype=torch.float32, device=dev), d)
    s.set_advector_generator(M.GC_DEFORM, 200.0, 0.4)
    torch.cuda.synchronize()
    s.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    const = torch.empty(shape, dtype=torch.float32, device=dev)
    s.get_advectee_device(const, n_tr - 1)
    dev_const = float((const - 1).abs().max())
    cells = shape[0] * shape[1] * shape[2]
    res = {"value": round(cells * n_tr * steps / el / 1e9, 4), "unit": "Gcell-updates/s",
           "ms_per_step": round(el * 1e3 / steps, 4), "steps": steps, "warmup": warmup,
           "config": {"workload": f"3D MPDATA-FCT n_iters=3, {shape[0]}x{shape[1]}x{shape[2]} cyclic, "
                                  f"{n_tr} tracers, reversing deformational advector generated each step",
                      "grid": list(shape), "tracers": n_tr},
           "alg_bytes_per_cell_step": 2 * n_tr * 4 + 3 * 4,
           "constant_tracer_max_dev": dev_const}
    s.close()
    return res


def cpu_baseline(seconds=10.0):
    """FP64 oracle (reference release-style flags, OpenMP) on a 128^3 sample
    of the same workload; FP64 like the reference's double-precision runs."""
    threads = int(os.environ.get("MPDX_CPU_THREADS", min(16, os.cpu_count() or 1)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle as O
    from libmpdataxx_amd import workloads as W
    n = 128
    psi, gc = W.rotation_sphere_3d(n)
    o = O.Oracle((n, n, n), opts=O.FCT, n_iters=3, fast=True)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.advance(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 5000:
            break
    return {"value": n ** 3 * steps / el / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n}^3 open, FCT n_iters=3, FP64, {steps} steps after 1 warm-up step, "
                      f"{el:.1f} s, OpenMP {threads} threads, -Ofast"}


def copy_bandwidth(torch, nbytes=1 << 30, reps=5):
    """Achievable HBM rate on this GPU: a device-to-device copy of nbytes,
    read + write bytes over the time of reps copies (HIP events on torch's
    stream, which the copies run on)."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1)
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes * reps / (ms * 1e-3) / 1e9


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--path", type=int, default=0, help="0 auto, 1 per-pass, 2 fused")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-multi", action="store_true", help="skip the multi-tracer leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the library steps one "
              "domain per GPU and has no multi-GPU halo exchange", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import libmpdataxx_amd as M

    n, K, W = args.grid, args.steps, args.warmup
    cells = n ** 3
    s32, el32, prof32, m0_32, elp32 = run_gpu(torch, M, n, "float32", K, W, args.path)
    value = cells * K / el32 / 1e9
    path_name = {1: "per-pass", 2: "fused"}.get(s32.active_path, "?")

    # dominant kernel: largest summed device time in the profiled steps
    dom = max(prof32.items(), key=lambda kv: kv[1][0])
    dname, (dms, dlaunch) = dom
    avg_ms = dms / max(dlaunch, 1)
    words, span = ALG_WORDS.get(dname, (0, "all"))
    launch_cells = (n - 4) ** 3 if span == "interior" else cells
    alg_bytes = words * 4 * launch_cells
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # HBM bytes per launch (FETCH_SIZE / WRITE_SIZE, copy-calibrated) and VALU
    # issue utilisation, recorded for the default 512^3 workload only
    default_wl = n == 512 and s32.active_path == 2
    traffic = profile_figure("pmc_traffic_fct3_512cube.md", "traffic_bytes_per_launch", dname) \
        if default_wl else None
    valu = profile_figure("pmc_sq_interior_fct3_512cube.md", "valu_issue_util", dname) if default_wl else None

    copy_gbs = copy_bandwidth(torch)
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(el32 * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (solid-body rotation of a discontinuous sphere, generated on device)",
        "config": {"workload": f"3D MPDATA-FCT n_iters=3, {n}^3, open bconds, FP32 state, FP64 guards",
                   "grid": [n, n, n], "n_iters": 3, "opts": "fct", "bcond": "open", "path": path_name},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(traffic) if traffic else None, "avg_launch_ms": round(avg_ms, 4),
                     "launches": dlaunch, "alg_bytes_per_launch": alg_bytes,
                     "valu_issue_util_pmc": valu, "copy_GBs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4)},
        "step_roofline": {"alg_bytes_per_cell": 20, "achieved_GBs": round(value * 20, 2),
                          "frac": round(value * 20 / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v[0], 3) for k, v in prof32.items()},
        "profiled_ms_per_step": round(elp32 * 1e3 / K, 4),
    }

    if not args.no_fp64:
        s64, el64, prof64, m0_64, _ = run_gpu(torch, M, n, "float64", K, W, args.path)
        v64 = cells * K / el64 / 1e9
        a32 = torch.empty((n, n, n), dtype=torch.float32, device="cuda:0")
        a64 = torch.empty((n, n, n), dtype=torch.float64, device="cuda:0")
        s32.get_advectee_device(a32)
        s64.get_advectee_device(a64)
        diff = a32.double() - a64
        l2 = float(torch.sqrt((diff ** 2).sum() / (a64 ** 2).sum()))
        linf = float(diff.abs().max())
        m32, m64 = s32.mass(), s64.mass()
        out["fp64"] = {"value": round(v64, 4), "ms_per_step": round(el64 * 1e3 / K, 4),
                       "kernels_ms": {k: round(v[0], 3) for k, v in prof64.items()}}
        out["fp32_over_fp64"] = round(value / v64, 4)
        out["accuracy_vs_fp64"] = {"l2_rel": l2, "linf": linf,
                                   "mass_rel_err_fp32": abs(m32 - m0_32) / abs(m0_32),
                                   "mass_rel_err_fp64": abs(m64 - m0_64) / abs(m0_64),
                                   "min_fp32": s32.min(), "steps": 2 * K + W}
        del a32, a64, diff
        s64.close()
    s32.close()

    if not args.no_multi:
        out["multi_tracer"] = run_multi_tracer(torch, M, max(K // 2, 2), min(W, 2))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

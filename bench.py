#!/usr/bin/env python3
"""Benchmark of the MPDATA advection step on one MI355X.

Headline workload: 3D MPDATA with FCT, n_iters = 3, open boundaries, 512^3
grid, FP32 state/advector/fluxes with FP64 guards and FP64 Kahan mass sums;
solid-body rotation about z (max Courant 0.5) moving a discontinuous sphere.
A "step" is one time step of the whole grid.  The same problem is also run
in FP64 for the FP32/FP64 ratio and as the accuracy reference (the FP64 GPU
path agrees with the FP64 CPU oracle to <= 1e-12), and in FP32 with the
guarded ratios evaluated in FP64.  Further legs: C2 (256^3 FP64), C1 (2D
256^2 FP64 cyclic), the multi-tracer workload (512x512x128 cyclic, 8
tracers, time-varying deformational advector generated on the device each
step), the uneven ensemble of 64 domains and the CPU baseline.  Inputs are resident in HBM before the timed
regions.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
METRIC = "Gcell-updates/s, 3D MPDATA-FCT iters=3, FP32 vs FP64, % of HBM peak"

# Algorithmic bytes per cell-update of one time step: every tracer read once
# and written once, every advector component read once; halo re-reads,
# intermediates (psi of the intermediate passes, antidiffusive velocities)
# and reductions excluded: prod(grid) x (2 T + D) x sizeof(real).
def alg_bytes_per_cell(tracers=1, dims=3, real_bytes=4):
    return (2 * tracers + dims) * real_bytes


def launch_cells(name, n):
    """Cells whose new value one launch of a kernel class computes, for the
    open-boundary n^3 workload: the interior kernels output the cells >= 2
    from the x and y boundaries and every z cell (their z-boundary tiles
    cover both open z boundaries); the other classes cover the whole grid."""
    if name.startswith("fused_first") or name.startswith("fused_interior"):
        return (n - 4) * (n - 4) * n
    return n ** 3


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


def run_gpu(torch, M, n, dtype, steps, warmup, path, guard=0, profile=True):
    dev = torch.device("cuda:0")
    s = M.Solver((n, n, n), opts=M.FCT, n_iters=3, dtype=dtype, path=path, guard=guard,
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
    if not profile:
        return s, (t1 - t0), {}, mass0, 0.0
    # the same number of steps again, launched eagerly with a HIP event pair
    # around every kernel on the solver's stream (per-kernel durations)
    s.profile(True)
    t2 = time.perf_counter()
    s.advance(steps)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    prof = s.profile_read()
    # the FP64 compensated mass sum and the max, timed the same way
    s.profile(True)
    for _ in range(5):
        s.mass()
        s.max()
    red = {k: v[0] / v[1] for k, v in s.profile_read().items()}
    s.profile(False)
    prof["_reductions_avg_ms"] = red
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
           "constant_tracer_max_dev": dev_const, "z_tiles": s.z_tiles}
    s.close()
    return res


def _cpu_leg(threads, seconds, runs, grid):
    """One child process timing the oracle's release-flag build with
    `threads` OpenMP threads (oracle/timing.py); returns its JSON."""
    import subprocess
    root = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-m", "oracle.timing", "--threads", str(threads), "--grid", str(grid),
                        "--seconds", str(seconds), "--runs", str(runs)],
                       cwd=root, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline run failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds=3.0, runs=3, grid=128):
    """The reference's CPU path as far as it can be run here: the FP64 oracle
    (same unfused per-pass structure and OpenMP y-slab threading as
    concurr::openmp, -Ofast, x86-64-v4 or -v3 to match the host) on a 128^3
    sample of the headline workload, timed with this process's CPU share of
    threads (OMP_NUM_THREADS, else every online CPU) and with one thread;
    median of `runs` timings of about `seconds` each, after a warm-up step."""
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    full = _cpu_leg(threads, seconds, runs, grid)
    one = _cpu_leg(1, seconds, runs, grid) if threads != 1 else full
    return {"value": round(full["gcell_per_s"], 6), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{grid}^3 open, FCT n_iters=3, FP64, median of {runs} runs of {full['steps_per_run']} steps "
                      f"after 1 warm-up step, OpenMP {threads} threads, -Ofast -march={full['isa']}",
            "one_thread": {"value": round(one["gcell_per_s"], 6), "steps_per_run": one["steps_per_run"],
                           "run_seconds": one["run_seconds"]},
            "run_seconds": full["run_seconds"], "nproc": nproc,
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "cpu_model": _cpu_model(), "isa": full["isa"],
            "note": "the reference's concurr::openmp itself needs Blitz++ and Boost, absent here; threads = the "
                    "process's CPU share (OMP_NUM_THREADS) rather than every host core"}


def run_ensemble(torch, M, steps, warmup, tune_steps=1):
    """C5: the uneven ensemble of 64 cyclic FP32 domains (32 2D from 64^2 to
    512^2, 32 3D from 32^3 to 128^3, n_iters 2-4, FCT on/off), autotuned,
    then stepped as one ensemble; beside it the same members advanced one
    after another with their own advance()."""
    import numpy as np
    from libmpdataxx_amd import workloads as W
    doms = W.ensemble_domains()
    mem = []
    for d in doms:
        nd = len(d["grid"])
        s = M.Solver(d["grid"], opts=M.FCT if d["fct"] else 0, n_iters=d["n_iters"], dtype=np.float32,
                     bcs=[(M.CYCLIC, M.CYCLIC)] * nd)
        s.set_advectee(d["psi"])
        for k in range(nd):
            s.set_advector(d["gc"][k], k)
        mem.append(s)
    cells = sum(int(np.prod(d["grid"])) for d in doms)
    e = M.Ensemble(mem)
    taken = e.tune(tune_steps)
    e.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.advance(steps)
    el = time.perf_counter() - t0
    launches, tasks = e.stats()
    tiles = sorted({(e.tile_y(i), e.x_chunk(i)) for i in range(len(mem)) if len(doms[i]["grid"]) == 3})
    e.close()
    for s in mem:
        s.advance(warmup)
    t0 = time.perf_counter()
    for s in mem:
        s.advance(steps)
    el_seq = time.perf_counter() - t0
    for s in mem:
        s.close()
    return {"value": round(cells * steps / el / 1e9, 4), "unit": "Gcell-updates/s",
            "ms_per_step": round(el * 1e3 / steps, 4), "steps": steps, "cells": cells, "dtype": "f32",
            "workload": "64 uneven cyclic domains (32 2D 64^2-512^2, 32 3D 32^3-128^3), n_iters 2-4, FCT on/off",
            "batched_launches_per_step": launches, "tasks_per_step": tasks, "tuning_steps": taken,
            "tuned_3d_tile_y_x_chunk": tiles,
            "members_one_by_one": {"value": round(cells * steps / el_seq / 1e9, 4),
                                   "ms_per_step": round(el_seq * 1e3 / steps, 4)}}


def run_2d_c1(torch, M, steps, warmup, n=256):
    """C1: 2D MPDATA n_iters=2, FP64, 256^2 cyclic, uniform Courant
    (0.25, 0.25), Gaussian cone on 0.5."""
    import numpy as np
    from libmpdataxx_amd import workloads as W
    psi, gc = W.gaussian_2d_cyclic(n)
    s = M.Solver((n, n), opts=0, n_iters=2, bcs=[(M.CYCLIC, M.CYCLIC)] * 2, dtype=np.float64)
    s.set_advectee(psi)
    for d in range(2):
        s.set_advector(gc[d], d)
    s.advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(steps)
    el = time.perf_counter() - t0
    s.close()
    return {"value": round(n * n * steps / el / 1e9, 4), "unit": "Gcell-updates/s",
            "ms_per_step": round(el * 1e3 / steps, 4), "steps": steps, "dtype": "f64",
            "workload": f"2D MPDATA n_iters=2, {n}^2 cyclic, uniform Courant (0.25, 0.25), Gaussian cone",
            "step_frac_of_hbm": round(n * n * steps / el * alg_bytes_per_cell(1, 2, 8) / 1e9 / HBM_PEAK_GBS, 5)}


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
    ap.add_argument("--cpu-seconds", type=float, default=3.0)
    ap.add_argument("--no-multi", action="store_true", help="skip the multi-tracer leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the uneven-ensemble leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the FP64-exact-guard, C1 (2D 256^2) and C2 (3D 256^3 FP64) legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the library steps one "
              "domain per GPU and has no multi-GPU halo exchange", file=sys.stderr)
        return 2

    import torch
    import libmpdataxx_amd as M

    n, K, W = args.grid, args.steps, args.warmup
    cells = n ** 3
    print(f"bench: {n}^3 FP32 FCT n_iters=3, {W} warm-up + {K} timed steps", file=sys.stderr, flush=True)
    s32, el32, prof32, m0_32, elp32 = run_gpu(torch, M, n, "float32", K, W, args.path)
    value = cells * K / el32 / 1e9
    path_name = {1: "per-pass", 2: "fused"}.get(s32.active_path, "?")
    bpc = alg_bytes_per_cell(1, 3, 4)          # 20 B per FP32 cell-update

    reductions = prof32.pop("_reductions_avg_ms")
    # dominant kernel: largest summed device time in the profiled steps
    dom = max(prof32.items(), key=lambda kv: kv[1][0])
    dname, (dms, dlaunch) = dom
    avg_ms = dms / max(dlaunch, 1)
    alg_bytes = bpc * launch_cells(dname, n)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) and VALU
    # issue utilisation, recorded by rocprofv3 for the default 512^3 workload
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
        "config": {"workload": f"3D MPDATA-FCT n_iters=3, {n}^3, open bconds, FP32 state and arithmetic, "
                               "FP64 guard constants, FP64 Kahan mass sums",
                   "grid": [n, n, n], "n_iters": 3, "opts": "fct", "bcond": "open", "path": path_name},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(traffic) if traffic else None,
                     "traffic_over_alg": round(traffic / alg_bytes, 3) if traffic else None,
                     "avg_launch_ms": round(avg_ms, 4), "launches": dlaunch,
                     "alg_bytes_per_cell": bpc, "cells_per_launch": launch_cells(dname, n),
                     "alg_bytes_per_launch": alg_bytes,
                     "valu_issue_util_pmc": valu, "copy_GBs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4)},
        "step_roofline": {"alg_bytes_per_cell": bpc, "achieved_GBs": round(value * bpc, 2),
                          "frac": round(value * bpc / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v[0], 3) for k, v in prof32.items()},
        "profiled_ms_per_step": round(elp32 * 1e3 / K, 4),
        "reduction_avg_ms": {k: round(v, 4) for k, v in reductions.items()},
    }

    if not args.no_fp64:
        print("bench: FP64 leg", file=sys.stderr, flush=True)
        s64, el64, prof64, m0_64, _ = run_gpu(torch, M, n, "float64", K, W, args.path)
        v64 = cells * K / el64 / 1e9
        a64 = torch.empty((n, n, n), dtype=torch.float64, device="cuda:0")
        s64.get_advectee_device(a64)

        def accuracy(sx, m0):
            a = torch.empty((n, n, n), dtype=torch.float32, device="cuda:0")
            sx.get_advectee_device(a)
            diff = a.double() - a64
            res = {"l2_rel": float(torch.sqrt((diff ** 2).sum() / (a64 ** 2).sum())),
                   "linf": float(diff.abs().max()),
                   "mass_rel_err": abs(sx.mass() - m0) / abs(m0), "min": sx.min(), "max": sx.max()}
            del a, diff
            return res
        red64 = prof64.pop("_reductions_avg_ms")
        out["fp64"] = {"value": round(v64, 4), "ms_per_step": round(el64 * 1e3 / K, 4),
                       "mass_rel_err": abs(s64.mass() - m0_64) / abs(m0_64),
                       "kernels_ms": {k: round(v[0], 3) for k, v in prof64.items()},
                       "reduction_avg_ms": {k: round(v, 4) for k, v in red64.items()}}
        out["fp32_over_fp64"] = round(value / v64, 4)
        out["accuracy_vs_fp64"] = dict(accuracy(s32, m0_32), steps=2 * K + W)
        s32.close()
        if not args.no_extra:
            # FP32 state with the guarded ratios evaluated in FP64 (IEEE FP64
            # division, FP64 constants), same steps as the FP64 run
            print("bench: FP32 FP64-exact-guard leg", file=sys.stderr, flush=True)
            sx, elx, _, m0_x, _ = run_gpu(torch, M, n, "float32", K, W, args.path, guard=M.GUARD_F64_EXACT,
                                          profile=False)
            vx = cells * K / elx / 1e9
            sx.advance(K)
            out["fp32_f64_exact_guards"] = {"value": round(vx, 4), "ms_per_step": round(elx * 1e3 / K, 4),
                                            "over_fp64": round(vx / v64, 4),
                                            "accuracy_vs_fp64": dict(accuracy(sx, m0_x), steps=2 * K + W)}
            sx.close()
        del a64
        s64.close()
    else:
        s32.close()
    torch.cuda.empty_cache()

    if not args.no_extra:
        print("bench: C2 256^3 FP64 leg", file=sys.stderr, flush=True)
        s2, el2, _, _, _ = run_gpu(torch, M, 256, "float64", K, W, args.path, profile=False)
        s2.close()
        out["c2_fp64_256"] = {"value": round(256 ** 3 * K / el2 / 1e9, 4), "unit": "Gcell-updates/s",
                              "ms_per_step": round(el2 * 1e3 / K, 4), "dtype": "f64",
                              "workload": "3D MPDATA-FCT n_iters=3, 256^3 open, solid-body rotation",
                              "step_frac_of_hbm": round(256 ** 3 * K / el2 * alg_bytes_per_cell(1, 3, 8) / 1e9
                                                        / HBM_PEAK_GBS, 4)}
        print("bench: C1 256^2 FP64 leg", file=sys.stderr, flush=True)
        out["c1_2d_fp64_256"] = run_2d_c1(torch, M, 10 * K, W)
    if not args.no_multi:
        print("bench: multi-tracer leg", file=sys.stderr, flush=True)
        out["multi_tracer"] = run_multi_tracer(torch, M, max(K // 2, 2), min(W, 2))
    if not args.no_ensemble:
        print("bench: ensemble leg", file=sys.stderr, flush=True)
        out["ensemble"] = run_ensemble(torch, M, K, W)
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Timed CPU baseline of the MPDATA step -- TEST INFRASTRUCTURE ONLY.

Runs the oracle's release-flag build (``-Ofast``, OpenMP, the x86-64 level
that matches the host: see ``fast_lib_name``) on the 3D MPDATA-FCT
solid-body-rotation workload and prints one JSON object.  Only ``bench.py``'s
``cpu_baseline`` leg runs it, as a child process, once per thread count (the
OpenMP team size is fixed when a process first enters a parallel region).

Timing is that of a libmpdata++ user timing concurr::openmp: one warm-up
step (the first advance runs hook_ante_loop: halos, divergence check), then
``std::chrono``-style wall clock around ``advance(k)``; the median of
``runs`` repetitions, with k chosen so that one repetition takes about
``seconds``.

    python -m oracle.timing --threads 16 --grid 128 --seconds 3 --runs 3
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, required=True)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--seconds", type=float, default=3.0)
    ap.add_argument("--runs", type=int, default=3)
    a = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    os.environ.setdefault("OMP_PROC_BIND", "false")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import oracle as O
    from libmpdataxx_amd import workloads as W

    n = a.grid
    psi, gc = W.rotation_sphere_3d(n)
    o = O.Oracle((n, n, n), opts=O.FCT, n_iters=3, fast=True)
    o.set_advectee(psi)
    for d in range(3):
        o.set_advector(gc[d], d)
    o.advance(1)                       # warm-up: hook_ante_loop and first step
    t0 = time.perf_counter()
    o.advance(1)
    one = max(time.perf_counter() - t0, 1e-6)
    k = max(1, int(round(a.seconds / one)))
    times = []
    for _ in range(a.runs):
        t0 = time.perf_counter()
        o.advance(k)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    print(json.dumps({"threads": a.threads, "grid": n, "steps_per_run": k, "runs": a.runs,
                      "run_seconds": [round(t, 4) for t in times], "median_s": med,
                      "gcell_per_s": n ** 3 * k / med / 1e9, "lib": O.fast_lib_name(),
                      "isa": O.host_isa()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""CPU baseline timer -- TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg).

Times the Yee step on the host cores on a bounded sample of the benchmark
workload: the 512^3 grid (or --n), the Gaussian-modulated Jz point source at
the centre, one warm-up step, then timed steps until --seconds have passed
(at least --min-steps, at most --max-steps).

kind "reference": the reference's own FDTD_openmp::FDTD (src/FDTD/FDTD.cpp)
compiled from its sources with its CMake flags (oracle/_ref); the source is
written to Jx = Jy = Jz before each step, as the reference drivers do
(perf-tests/sample/sample.cpp:67-82), so the OpenMP class injects it.
kind "port": the oracle's C restatement (same loop structure and flags).
With --pml P the grid is the PML class (src/FDTD/FDTD_PML.cpp) with
pml_percent P, resp. the oracle's restatement of it; the reference class
runs on the whole OpenMP team as its driver does (sample.cpp:93-123).

Run as a subprocess so that OMP_NUM_THREADS takes effect:
    OMP_NUM_THREADS=16 python -m oracle.cpu_bench --n 512
Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seconds", type=float, default=12.0)
    ap.add_argument("--min-steps", type=int, default=3)
    ap.add_argument("--max-steps", type=int, default=25)
    ap.add_argument("--kind", choices=["auto", "reference", "port"], default="auto")
    ap.add_argument("--pml", type=float, default=None)
    a = ap.parse_args()
    kind = a.kind
    if kind == "auto":
        kind = "reference" if O.ref_available() else "port"
    n = a.n
    d = 1.0
    dt = 0.5 * d / (O.C * np.sqrt(3.0))
    src = dict(i=n // 2, j=n // 2, k=n // 2, comp=2, kind=1, amp=1.0, period=20 * dt,
               t0=120 * dt, sigma=40 * dt, phase=0.0)
    idx = src["i"] + src["j"] * n + src["k"] * n * n
    t_setup = time.perf_counter()
    if kind == "reference":
        if a.pml is None:
            g = O.RefGrid(n, n, n, d, d, d, dt)
            advance = g.update_fields
        else:
            g = O.RefPmlGrid(n, n, n, d, d, d, dt, a.pml)

            def advance(k):
                g.update_fields(k, one_thread=False)

        def step(s):
            val = O.waveform(1, 1.0, src["period"], src["t0"], src["sigma"], 0.0, (s + 1) * dt)
            for comp in (6, 7, 8):
                g.set_value(comp, idx, val)
            advance(1)
    else:
        g = (O.OracleGrid(n, n, n, d, d, d, dt) if a.pml is None
             else O.OraclePmlGrid(n, n, n, d, d, d, dt, a.pml))

        def step(s):
            g.run_point_sources([src], 1)
    t_setup = time.perf_counter() - t_setup
    step(0)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        step(steps + 1)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= a.max_steps or (steps >= a.min_steps and el >= a.seconds):
            break
    cells = float(n) ** 3
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    print(json.dumps({
        "value": cells * steps / el / 1e9, "unit": "Gcell-updates/s", "cores": threads,
        "kind": kind, "steps": steps, "seconds": el, "setup_seconds": t_setup,
        "sample": (f"{n}^3 " + ("periodic grid" if a.pml is None else
                                f"grid with a PML layer (pml_percent {a.pml})")
                   + ", Gaussian-modulated Jz point source at the centre, "
                   f"1 warm-up + {steps} timed steps (B/2,E,B/2 each), "
                   f"OMP_NUM_THREADS={threads}"),
    }))


if __name__ == "__main__":
    main()

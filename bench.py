#!/usr/bin/env python3
"""Benchmark of the Yee time step on one MI355X.

Workload (the headline configuration): a 512^3 periodic vacuum grid,
dx = dy = dz = 1, dt = 0.5 dx / (c sqrt 3), a Gaussian-modulated sinusoid
Jz point source at the centre (period 20 dt, envelope sigma 40 dt, peak at
120 dt, amplitude 1, sampled at t = (n+1) dt), fields starting at zero,
K steps (default 1000).  A step is one reference update_fields()
(B/2, E, B/2) over the whole grid.

Each precision runs K steps with the full-row fused sweep, then K steps
with the configuration the blocking autotuner picks (the fused sweep,
ghost-zone patch passes or wavefront-strip passes of L steps each; all
bit-identical to single sweeps).

Reported:
  value         FP32, autotuned path, Gcell-updates/s (wall clock over K
                steps, synchronised on both sides)
  roofline      the steady kernel of that path: achieved = one read and
                one write of E and B (48 B per cell in FP32) per launch /
                the average launch time (HIP events around each launch on
                the grid's stream, inside the timed region), frac = that /
                8 TB/s; traffic = measured HBM bytes per launch (rocprofv3
                FETCH_SIZE/WRITE_SIZE runs committed in
                profiles/hbm_traffic.json) and traffic_frac its rate;
                valu_busy from the same counter runs; for blocked passes
                effective_frac restates the pass as L single steps
  fused_sweep_roofline  the same for the single-step fused sweep
  paths         FP32/FP64/FP16 autotuned and fused, and the
                reference-structured FP64 three-sweep path
  errors_vs_fp64  relative L2 and max-abs error of FP32 and FP16 after K
                steps against the FP64 run, per component
  cpu_baseline  the reference's own OpenMP class (compiled from its
                sources) on this host's cores, bounded sample

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
python bench.py --profile fp32|fp64|fp16 [--tuning V,TY,kchunk] [--blocking L,V,R,W]
    (only that path: for rocprofv3)

Other workloads (not the headline; each prints one JSON line):
python bench.py --workload batch [--grids 256] [--steps 500] [--blocking L,V,R,W]
    256 independent FP32 grids, each dimension from {32,48,...,160} (seeded),
    one plane wave (wavelength 8-32 cells) and one Gaussian-modulated Jz
    source each: one persistent launch per step, HIP-graph captured,
    autotuned shape; aggregate Gcell/s, and ms/step against one launch per grid
python bench.py --workload blocked [--dims 1024,1024,256] [--steps 4000]
    FP32 with 4 seeded Jz sources of distinct periods: the full-row sweep
    against the autotuned best of tiled/temporally blocked passes
    (L steps per pass over HBM); effective roofline, error vs FP64
python bench.py --workload pml [--n 512] [--pml 0.2] [--steps 1000]
    the centre-source workload on the PML class (absorbing layer of
    int(N*pml) cells per side, as the reference driver's PML block): FP32 and
    FP64, each with the two in-place phase sweeps per step and with the fused
    PML sweep; roofline over 12 words per main-box cell and 36 per layer cell
    per step, error vs FP64, CPU baseline of the reference PML class (256^3
    sample)
python bench.py --workload precision [--n 768] [--steps 2000]
    768^3, superposition of 8 seeded plane waves plus the centre Jz source:
    FP64, FP32 and FP16 (per-component power-of-two scales) runs, per
    component error and energy drift against FP64
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fdtd_method_amd import FDTD, Parameters, Precision, calibration_copy  # noqa: E402
from fdtd_method_amd import _lib  # noqa: E402

C = 3e10
PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + %HBM roofline, FP32/FP16 vs FP64 GPU at 512³, L2 err vs FP64"
WORD = {Precision.FP64: 8, Precision.FP32: 4, Precision.FP16: 2}
PML_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_pml_sweep_fp32_512_pml0.2_traffic.json")
PROFILE_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def committed_traffic(path, **match):
    """HBM bytes per launch from a committed counter summary under
    profiles/ whose workload/configuration matches (a file may hold several
    configurations under "configs"); None if there is none."""
    if not os.path.exists(path):
        return None
    t = json.load(open(path))
    for rec in t.get("configs", [t]):
        if all(rec.get(k) == v for k, v in match.items()):
            return rec.get("hbm_bytes_per_launch")
    return None


def committed_profile(prec, n, blocking, tuning):
    """The committed counter record (profiles/hbm_traffic.json) of the steady
    kernel of a 512^3-workload path: matched by precision, grid, and the
    blocking (L, V, R, W, kchunk and stored tile width) or the fused launch
    shape (V, TY, kchunk) the run used; None when no counter run of that
    exact configuration is committed."""
    if not os.path.exists(PROFILE_FILE):
        return None
    want = {"precision": prec.name.lower(), "n": n}
    if blocking:
        want["kernel"] = "strip" if blocking.get("R", 0) > 0 else "patch"
        want["blocking"] = [blocking["L"], blocking["V"], blocking.get("R", 0),
                            blocking.get("W", 0), blocking["kchunk"]]
    else:
        want["kernel"] = "fused"
        want["tuning"] = list(tuning)
    for rec in json.load(open(PROFILE_FILE))["configs"]:
        if all(rec.get(k) == v for k, v in want.items()):
            return dict(rec, file=os.path.relpath(PROFILE_FILE, ROOT))
    return None


def make_grid(n, precision, autotune=True):
    dt = 0.5 / (C * np.sqrt(3.0))
    g = FDTD(Parameters(n, n, n), dt, precision)
    g.set_time_blocking_strips(0)  # single sweeps until the blocking autotuner runs
    src = dict(i=n // 2, j=n // 2, k=n // 2, comp=2, kind=1, amp=1.0, period=20 * dt,
               t0=120 * dt, sigma=40 * dt, phase=0.0)
    g.set_point_sources([src])
    if precision == Precision.FP16:
        # fields start at zero: seed the scales from one step's injection
        # |cur * J| ~ 4 pi dt, then rescale the running state every 20 steps
        g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * dt * 1e-2)
        g.set_autoscale_interval(20, 10)
    if autotune:
        shape, ms = g.autotune(reps=4)
        log(f"autotune {precision.name}: V={shape[0]} TY={shape[1]} kchunk={shape[2]} "
            f"({ms:.3f} ms/sweep)")
    return g


def reset(g):
    """Zero fields, step 0; FP16 scales sized for one step's injection."""
    g.zero_fields()
    if g.precision == Precision.FP16:
        g.set_scales([1.0] * 6)
        g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * g.dt * 1e-2)


def run_path(g, steps, warmup, path=_lib.PATH_FUSED):
    """From zero fields: `warmup` untimed steps, then `steps` timed steps that
    continue the same run (so the timed call starts from the half step the
    warm-up's last launch carried, as a driver's next call would).  Wall
    clock around the timed call, and HIP events around each of its
    launches on the grid's stream (fused path) or around the whole call
    (three-sweep path)."""
    reset(g)
    if warmup:
        g.time_update_fields(warmup, path)
    g.synchronize()
    cells = float(g.cells)
    t0 = time.perf_counter()
    if path == _lib.PATH_FUSED:
        kern = g.time_update_fields_kernels(steps)
        g.synchronize()
        wall = time.perf_counter() - t0
        main = sum(v[0] for v in kern.values())
        launches = sum(v[1] for v in kern.values())
    else:
        _, main = g.time_update_fields(steps, path)
        g.synchronize()
        wall = time.perf_counter() - t0
        kern, launches = None, steps
    rec = dict(ms_per_step=1e3 * wall / steps, gcells=cells * steps / wall / 1e9,
               event_main_ms=main, launch_ms=main / max(1, launches), kernels=kern)
    if kern:
        # the steady launch: the kernel a long run is made of
        k = kern["steady"] if kern["steady"][1] else kern["last"]
        rec["steady_launch_ms"] = k[0] / k[1]
        if kern["last"][1]:
            rec["last_launch_ms"] = kern["last"][0] / kern["last"][1]
    return rec


def errors(test, ref):
    per = test.compare(ref)
    names = ["Ex", "Ey", "Ez", "Bx", "By", "Bz"]
    # Per-component errors are taken relative to the norm of all six
    # reference components: a component that stays zero up to rounding (Bz
    # under Jz-only sources) has no meaningful norm of its own.
    comp = {}
    sr_all = sum(p[1] for p in per)
    for nm, (sd, sr, mx) in zip(names, per):
        comp[nm] = dict(rel_l2_of_total=float(np.sqrt(sd / sr_all)) if sr_all > 0 else None,
                        max_abs=mx)
    tot = float(np.sqrt(sum(p[0] for p in per) / sr_all))
    scale = max(float(np.sqrt(p[1] / test.cells)) for p in per)
    return dict(rel_l2=tot, max_abs=max(p[2] for p in per), rms_ref=scale, per_component=comp)


def energy(g):
    s = g.sum_squares()
    return sum(s) / (8.0 * np.pi)


# Host threads for the CPU baseline: one per physical core available to this
# process, at most 16 -- the host share of one GPU on the machines this runs
# on (a box of one GPU grants its process 16 hardware threads).
CPU_SHARE = 16


def host_cpu():
    """Model, sockets, physical cores and hardware threads of the host
    (/proc/cpuinfo), and how many of them this process may use."""
    info = dict(model=None, sockets=None, physical_cores=None, hw_threads=os.cpu_count())
    cores, sockets, model = {}, set(), None
    try:
        proc = phys = core = None
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "processor":
                proc = int(val)
            elif key == "physical id":
                phys = val
            elif key == "core id":
                core = val
            elif key == "model name" and model is None:
                model = val
            elif not key and proc is not None:
                cores[proc] = (phys, core)
                sockets.add(phys)
                proc = phys = core = None
        if proc is not None:
            cores[proc] = (phys, core)
            sockets.add(phys)
    except OSError:
        pass
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = list(range(os.cpu_count() or 1))
    info.update(model=model, sockets=len(sockets) or None,
                physical_cores=len(set(cores.values())) or None,
                allowed_hw_threads=len(allowed),
                allowed_physical_cores=len({cores[p] for p in allowed if p in cores}) or None)
    return info


def cpu_baseline(n, pml=None):
    host = host_cpu()
    threads = min(CPU_SHARE, host.get("allowed_physical_cores") or host["allowed_hw_threads"])
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    cmd = [sys.executable, "-m", "oracle.cpu_bench", "--n", str(n), "--seconds", "12"]
    if pml is not None:
        cmd += ["--pml", str(pml)]
    log("cpu baseline:", " ".join(cmd), f"(OMP_NUM_THREADS={threads})")
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT, env=env, timeout=600)
    if res.returncode != 0:
        log(res.stderr[-2000:])
        return None
    rec = json.loads(res.stdout.strip().splitlines()[-1])
    out = {k: rec[k] for k in ("value", "unit", "cores", "kind", "sample")}
    out["host"] = host
    out["threads_rule"] = (f"one OpenMP thread per physical core available, at most {CPU_SHARE} "
                           "(the host share of one GPU), OMP_PROC_BIND=close, OMP_PLACES=cores")
    return out


def bench_batch(a):
    from fdtd_method_amd import Batch, plane_wave_batch
    K = a.steps
    t0 = time.perf_counter()
    grids, desc = plane_wave_batch(n_grids=a.grids, seed=a.seed)
    log(f"batch: {len(grids)} grids, {sum(g.cells for g in grids) / 1e6:.1f} M cells, "
        f"built in {time.perf_counter() - t0:.1f} s")
    b = Batch(grids)
    shape, ms_tune = b.autotune(reps=3)
    log(f"batch autotune: {shape} ({ms_tune:.3f} ms/sweep)")
    cells = float(b.cells)

    def reset():
        for q, (g, d) in enumerate(zip(grids, desc)):
            g.zero_fields()
            g.add_plane_wave(d["m"], d["amp"], d["phase"],
                             np.random.default_rng(a.seed * 1000 + q).standard_normal(3))

    def parity():
        # batched graph replays == single-grid runs, on the first grids
        reset()
        ref, _ = plane_wave_batch(n_grids=min(8, a.grids), seed=a.seed)
        b.update_fields_graph(12, 2)
        for g in ref:
            g.update_fields(24)
        same = all(r.checksum() == g.checksum() for r, g in zip(ref, grids))
        for g in ref:
            g.close()
        return same

    def timed_run():
        reset()
        b.update_fields_graph(K, 0)  # capture only
        t0 = time.perf_counter()
        b.update_fields_graph(K, 1)
        return time.perf_counter() - t0

    # one batched fused sweep per step
    same_fused = parity()
    wall_fused = timed_run()
    ms_b, ms_s = b.time(min(K, 48))
    # persistent strip passes of L steps over every grid's tiles
    best = None
    cands = ((4, 1, 2, 16), (4, 1, 2, 8), (3, 1, 2, 8), (2, 1, 2, 8), (4, 1, 4, 4))
    if a.blocking:
        cands = (tuple(int(x) for x in a.blocking.split(","))[:4],)
    for cfg in cands:
        try:
            b.set_time_blocking_strips(*cfg)
        except RuntimeError as e:
            log(f"batch strips {cfg}: {e}")
            continue
        ms, _ = b.time(48, sequential=False)
        log(f"batch strips {cfg}: {ms:.4f} ms/step ({b.get_time_blocking_strips()['work_items']} "
            "tiles)")
        if best is None or ms < best[1]:
            best = (cfg, ms)
    rec_strips = None
    wall = wall_fused
    if best is not None:
        b.set_time_blocking_strips(*best[0])
        same_strips = parity()
        wall_strips = timed_run()
        rec_strips = {"blocking_L_V_R_W": list(best[0]),
                      "ms_per_step": 1e3 * wall_strips / K,
                      "gcells_per_s": cells * K / wall_strips / 1e9,
                      "speedup_vs_fused_batch": wall_fused / wall_strips,
                      "bit_identical_to_single_grid_runs": same_strips}
        if same_strips and wall_strips < wall_fused:
            wall = wall_strips
    path = "strips" if wall is not wall_fused else "fused"
    out = {
        "metric": "aggregate Gcell-updates/s, batched persistent launches",
        "value": cells * K / wall / 1e9, "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K,
        "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{len(grids)} independent periodic FP32 grids, dims from "
                                "{32,48,...,160} (seed %d), one plane wave and one "
                                "Gaussian-modulated Jz each, %d steps, graph-captured" % (a.seed, K)),
                   "cells": cells, "shape": shape, "path": path},
        "fused": {"ms_per_step": 1e3 * wall_fused / K,
                  "gcells_per_s": cells * K / wall_fused / 1e9,
                  "alg_gbs": 48.0 * cells / (wall_fused / K) / 1e9,
                  "batched_ms_per_step": ms_b, "sequential_ms_per_step": ms_s,
                  "speedup_vs_one_launch_per_grid": ms_s / ms_b,
                  "bit_identical_to_single_grid_runs": same_fused},
        "strips": rec_strips,
    }
    print(json.dumps(out), flush=True)


def bench_precision(a):
    n, K = a.n, a.steps
    dt = 0.5 / (C * np.sqrt(3.0))
    rng = np.random.default_rng(a.seed)
    waves = []
    while len(waves) < 8:
        m = [int(x) for x in rng.integers(-8, 9, 3)]
        if not any(m):
            continue
        waves.append(dict(m=m, amp=float(rng.uniform(0.1, 1.0)),
                          phase=float(rng.uniform(0, 2 * np.pi)),
                          pol=[float(x) for x in rng.standard_normal(3)]))
    src = dict(i=n // 2, j=n // 2, k=n // 2, comp=2, kind=1, amp=1.0, period=20 * dt,
               t0=120 * dt, sigma=40 * dt)
    res, grids = {}, {}
    for prec in (Precision.FP64, Precision.FP32, Precision.FP16):
        g = FDTD(Parameters(n, n, n), dt, prec)
        if prec == Precision.FP16:
            g.set_scales([2.0 ** 10] * 6)  # sum of 8 amplitudes <= 8 -> max ~2^13
        for w in waves:
            g.add_plane_wave(w["m"], w["amp"], w["phase"], w["pol"])
        if prec == Precision.FP16:
            g.autoscale(target_log2=12)
            g.set_autoscale_interval(100, 12)
        g.set_point_sources([src])
        g.autotune(reps=3)
        blocking, _ = g.autotune_blocking(reps=4)  # bit-identical either way
        e0 = energy(g)
        g.synchronize()
        t0 = time.perf_counter()
        ms_total, ms_main = g.time_update_fields(K)
        wall = time.perf_counter() - t0
        res[prec.name] = dict(gcells=float(n) ** 3 * K / wall / 1e9, ms_per_step=1e3 * wall / K,
                              energy_start=e0, energy_end=energy(g), blocking=blocking)
        grids[prec.name] = g
        log(f"{prec.name}: {res[prec.name]['gcells']:.1f} Gcell/s, blocking {blocking}")
    e32 = errors(grids["FP32"], grids["FP64"])
    e16 = errors(grids["FP16"], grids["FP64"])
    e64 = res["FP64"]["energy_end"]
    for nm in ("FP32", "FP16"):
        res[nm]["energy_drift_vs_fp64"] = res[nm]["energy_end"] / e64 - 1.0
    out = {
        "metric": "relative L2 / max-abs error vs FP64 and Gcell-updates/s per precision",
        "value": e32["rel_l2"], "unit": "relative L2 (FP32 vs FP64)", "n_gpus": 1, "steps": K,
        "higher_is_better": False, "dtype": "f32/f16 vs f64", "data": "synthetic",
        "config": {"workload": (f"{n}^3 periodic grid, 8 seeded plane waves (seed {a.seed}) "
                                f"plus the centre Gaussian-modulated Jz, {K} steps"),
                   "waves": waves},
        "paths": res, "errors_vs_fp64": {"fp32": e32, "fp16": e16},
    }
    print(json.dumps(out), flush=True)


def bench_blocked(a):
    Ni, Nj, Nk = (int(x) for x in a.dims.split(","))
    K = a.steps
    dt = 0.5 / (C * np.sqrt(3.0))
    rng = np.random.default_rng(a.seed)
    periods = rng.permutation(np.arange(16, 41))[:4]
    srcs = [dict(i=int(rng.integers(Ni)), j=int(rng.integers(Nj)), k=int(rng.integers(Nk)),
                 comp=2, kind=1, amp=1.0, period=float(p) * dt, t0=0.0, sigma=0.0)
            for p in periods]
    cells = float(Ni) * Nj * Nk

    def make(prec):
        g = FDTD(Parameters(Ni, Nj, Nk), dt, prec)
        g.set_time_blocking_strips(0)  # the full-row sweep unless set below
        g.set_point_sources(srcs)
        return g

    g = make(Precision.FP32)
    g.autotune(reps=3)
    full_shape = g.get_tuning()
    # full-row fused sweep over the whole run
    g.time_update_fields(4)
    g.zero_fields()
    t0 = time.perf_counter()
    _, ms_full = g.time_update_fields(K)
    wall_full = time.perf_counter() - t0
    ms_full /= K
    # blocked passes: short timing of candidate configurations, then the run
    g.zero_fields()
    best = None
    # wavefront-strip kernel {L, V, R, W} and ghost-zone patch kernel {L, V, lanes, TYg}
    cands = [("strips", (L, V, R, W)) for L in (4, 3, 2)
             for V, R, W in ((2, 2, 8), (1, 2, 16), (1, 2, 8), (1, 4, 8))]
    cands += [("patch", c) for c in ((4, 1, 32, 32), (3, 1, 32, 32), (3, 1, 64, 16))]

    def apply(kind, c):
        if kind == "strips":
            g.set_time_blocking_strips(*c, 0)
        else:
            g.set_time_blocking(*c, 0)

    def describe(kind):
        return g.get_time_blocking_strips() if kind == "strips" else g.get_time_blocking()

    for kind, cand in cands:
        try:
            apply(kind, cand)
        except Exception:
            continue
        L = cand[0]
        g.time_update_fields(2 * L)
        _, m = g.time_update_fields(12 * L)
        m /= 12 * L
        log(f"  {kind} {describe(kind)}: {m:.3f} ms/step")
        if best is None or m < best[0]:
            best = (m, kind, cand)
    apply(best[1], best[2])
    cfg = dict(kernel=best[1], **describe(best[1]))
    g.zero_fields()
    g.synchronize()
    t0 = time.perf_counter()
    _, ms_main = g.time_update_fields(K)
    wall = time.perf_counter() - t0
    log(f"full-row {ms_full:.3f} ms/step, blocked {cfg} {ms_main / K:.3f} ms/step")
    ref = make(Precision.FP64)
    ref.autotune(reps=2)
    ref.time_update_fields(K)
    err = errors(g, ref)
    eff = 48.0 * cells / (ms_main / K * 1e-3) / 1e9
    out = {
        "metric": "Gcell-updates/s and effective %HBM roofline with temporal blocking",
        "value": cells * K / wall / 1e9, "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K,
        "ms_per_step": 1e3 * wall / K, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{Ni}x{Nj}x{Nk} periodic FP32 grid, 4 Jz point sources "
                                f"(periods {[int(p) for p in periods]} dt, seed {a.seed}), {K} steps"),
                   "blocking": cfg, "full_row_shape": full_shape},
        "effective_gbs": eff, "effective_frac_of_8tbs": eff / PEAK_GBS,
        "full_row_ms_per_step": ms_full, "blocked_ms_per_step": ms_main / K,
        "full_row_gcells": cells * K / wall_full / 1e9,
        "full_row_effective_frac_of_8tbs": 48.0 * cells / (ms_full * 1e-3) / 1e9 / PEAK_GBS,
        "speedup_vs_full_row": ms_full / (ms_main / K),
        "errors_vs_fp64": err,
    }
    print(json.dumps(out), flush=True)


def pml_bytes_per_step(shape, widths, word):
    """Algorithmic HBM bytes of one PML step: main-box cells read and write
    E and B (12 words); layer cells also read and write their 12 split
    fields (36 words)."""
    Ni, Nj, Nk = shape
    main = 1
    for n, p in zip(shape, widths):
        main *= max(0, n - 2 * p)
    layer = Ni * Nj * Nk - main
    return word * (12 * main + 36 * layer), main, layer


def bench_pml(a):
    """The PML workload on both PML step forms: the two in-place phase sweeps
    per step (pml_e_phase, pml_b_phase) and the fused PML sweep.  `value` is
    the grid's default form in FP32."""
    from fdtd_method_amd import FDTD_PML
    n, K, pml = a.n, a.steps, a.pml
    dt = 0.5 / (C * np.sqrt(3.0))
    src = dict(i=n // 2, j=n // 2, k=n // 2, comp=2, kind=1, amp=1.0, period=20 * dt,
               t0=120 * dt, sigma=40 * dt, phase=0.0)
    cells = float(n) ** 3

    def make(prec, form):
        g = FDTD_PML(Parameters(n, n, n), dt, pml, prec)
        g.set_point_sources([src])
        if form == "phases":
            g.set_pml_path(1)
        if form == "fused":
            g.set_pml_path(0)
            if a.tuning:
                g.set_tuning(*[int(x) for x in a.tuning.split(",")])
            else:
                shape, ms = g.autotune(reps=3)
                log(f"pml autotune {prec.name}: V={shape[0]} TY={shape[1]} kchunk={shape[2]} "
                    f"({ms:.3f} ms/sweep)")
        return g

    default_form = "phases" if make(Precision.FP32, "default").get_pml_path()["path"] == 1 \
        else "fused"
    if a.profile:
        # one precision only, for rocprofv3 runs
        prec = {"fp32": Precision.FP32, "fp64": Precision.FP64, "fp16": Precision.FP16}[a.profile]
        g = make(prec, a.pml_form or default_form)
        r = run_path(g, K, a.warmup)
        nbytes, main, layer = pml_bytes_per_step((n, n, n), g.pml_widths, WORD[prec])
        ms_copy = calibration_copy(1 << 30, 5)
        print(json.dumps(dict(profile=a.profile, workload="pml", n=n, pml=pml, steps=K,
                              form=a.pml_form or default_form, pml_path=g.get_pml_path(),
                              tuning=g.get_tuning(), widths=g.pml_widths,
                              alg_bytes_per_step=nbytes, main_cells=main, layer_cells=layer,
                              **r, calibration_copy=dict(bytes=1 << 30, ms=ms_copy))),
              flush=True)
        return

    res = {}
    grids = {}
    for prec in (Precision.FP32, Precision.FP64):
        for form in ("phases", "fused"):
            g = make(prec, form)
            r = run_path(g, K, a.warmup)
            nbytes, main, layer = pml_bytes_per_step((n, n, n), g.pml_widths, WORD[prec])
            gbs = nbytes / (r["steady_launch_ms"] * 1e-3) / 1e9
            rec = dict(gcells=r["gcells"], ms_per_step=r["ms_per_step"],
                       step_kernel_ms=r["steady_launch_ms"], alg_bytes_per_step=nbytes,
                       alg_gbs=gbs, frac_of_peak=gbs / PEAK_GBS)
            if form == "fused":
                rec["launch_shape"] = g.get_tuning()
            else:
                rec["phase_shape"] = g.get_pml_path()
                # the phases read E and B of main cells once more per step
                wd = WORD[prec]
                rec["phase_bytes_per_step"] = wd * (18 * main + 36 * layer)
            res[f"{prec.name.lower()}_{form}"] = rec
            log(f"pml {prec.name} {form}: {r['gcells']:.1f} Gcell/s, step kernels "
                f"{r['steady_launch_ms']:.3f} ms, {gbs:.0f} GB/s algorithmic")
            if form == default_form:
                grids[prec] = g
            else:
                g.close()
    err = errors(grids[Precision.FP32], grids[Precision.FP64])
    widths = grids[Precision.FP32].pml_widths
    en = energy(grids[Precision.FP64])
    nbytes, main, layer = pml_bytes_per_step((n, n, n), widths, 4)
    r32 = res[f"fp32_{default_form}"]
    kern = ("fdtd::pml_e_phase<float,float,float,V> + fdtd::pml_b_phase<float,float,float,V,true>"
            if default_form == "phases" else "fdtd::fused_pml_sweep<float,float,float,V,true>")
    cpu = None if a.no_cpu_baseline else cpu_baseline(min(n, 256), pml)
    out = {
        "metric": "Gcell-updates/s and %HBM roofline of the PML (absorbing layer) step",
        "value": r32["gcells"], "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K,
        "warmup": a.warmup, "ms_per_step": r32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"{n}^3 grid with a PML layer (pml_percent {pml}: widths "
                                f"{list(widths)}, {layer / cells:.1%} of the cells), dx=1, "
                                "dt=0.5dx/(c*sqrt3), Gaussian-modulated Jz point source at the "
                                f"centre, zero initial fields, {K} steps"),
                   "main_box_cells": main, "layer_cells": layer, "pml_step_form": default_form},
        "roofline": {"bound": "hbm", "achieved": r32["alg_gbs"], "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": r32["alg_gbs"] / PEAK_GBS,
                     "traffic": committed_traffic(PML_TRAFFIC_FILE, n=n, pml=pml,
                                                  form=default_form,
                                                  tuning=list(grids[Precision.FP32].get_tuning())),
                     "kernel": kern, "alg_bytes_per_step": nbytes,
                     "step_kernel_ms": r32["step_kernel_ms"]},
        "cpu_baseline": cpu,
        "paths": res,
        "errors_vs_fp64": {"fp32": err},
        "energy_fp64": en,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="steps (default: point 1000, batch 500, precision 2000, blocked 4000, "
                         "pml 1000)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=None,
                    help="grid edge (default: precision 768, otherwise 512)")
    ap.add_argument("--three-sweep-steps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", choices=["fp32", "fp64", "fp16"], default=None)
    ap.add_argument("--tuning", default=None,
                    help="V,TY,kchunk launch shape for --profile (default: autotune)")
    ap.add_argument("--blocking", default=None,
                    help="L,V,R,W[,kchunk] strip blocking for --profile (default: none); "
                         "L,V,R,W for --workload batch (default: the best of five)")
    ap.add_argument("--workload", choices=["point", "batch", "precision", "blocked", "pml"],
                    default="point")
    ap.add_argument("--pml", type=float, default=0.2, help="pml_percent of --workload pml")
    ap.add_argument("--pml-form", choices=["phases", "fused"], default=None,
                    help="PML step form for --workload pml --profile (default: the grid's)")
    ap.add_argument("--dims", default="1024,1024,256", help="grid of --workload blocked")
    ap.add_argument("--grids", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): there is no "
              "multi-GPU path (the grid would shard into z-slabs with halo-plane exchange, "
              "not implemented).", file=sys.stderr)
        sys.exit(2)
    if a.steps is None:
        a.steps = {"batch": 500, "precision": 2000, "blocked": 4000}.get(a.workload, 1000)
    if a.n is None:
        a.n = 768 if a.workload == "precision" else 512
    n, K, W = a.n, a.steps, a.warmup
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "precision":
        return bench_precision(a)
    if a.workload == "blocked":
        return bench_blocked(a)
    if a.workload == "pml":
        return bench_pml(a)

    if a.profile:
        prec = {"fp32": Precision.FP32, "fp64": Precision.FP64, "fp16": Precision.FP16}[a.profile]
        g = make_grid(n, prec, autotune=a.tuning is None)
        if a.tuning:
            g.set_tuning(*[int(x) for x in a.tuning.split(",")])
        blk = None
        if a.blocking:
            g.set_time_blocking_strips(*[int(x) for x in a.blocking.split(",")])
            blk = g.get_time_blocking_strips()
        r = run_path(g, K, W)
        nbytes = 1 << 30
        ms_copy = calibration_copy(nbytes, 5)
        print(json.dumps(dict(profile=a.profile, n=n, steps=K, tuning=g.get_tuning(),
                              blocking=blk, **r,
                              calibration_copy=dict(bytes=nbytes, ms=ms_copy,
                                                    gbs=2 * nbytes / ms_copy / 1e6))),
              flush=True)
        return

    log(f"512^3 workload: n={n} steps={K} warmup={W}")
    cells = float(n) ** 3
    grids, fused, best, blocking = {}, {}, {}, {}
    for prec in (Precision.FP32, Precision.FP64, Precision.FP16):
        g = make_grid(n, prec)
        fused[prec] = run_path(g, K, W)
        log(f"{prec.name} fused: {fused[prec]['gcells']:.1f} Gcell/s, "
            f"launch {fused[prec]['launch_ms']:.3f} ms")
        cfg, ms = g.autotune_blocking(reps=8)
        blocking[prec] = cfg
        if cfg["L"] > 0:
            best[prec] = run_path(g, K, W)
            log(f"{prec.name} blocked {cfg}: {best[prec]['gcells']:.1f} Gcell/s")
        else:
            best[prec] = fused[prec]
        grids[prec] = g
    g32, g64, g16 = grids[Precision.FP32], grids[Precision.FP64], grids[Precision.FP16]
    e32 = errors(g32, g64)
    e16 = errors(g16, g64)
    en = dict(fp64=energy(g64), fp32=energy(g32), fp16=energy(g16))
    log(f"errors vs fp64: fp32 relL2={e32['rel_l2']:.3e}  fp16 relL2={e16['rel_l2']:.3e}")
    tunes = {p: grids[p].get_tuning() for p in grids}
    g16.close()
    g32.close()
    # reference-structured FP64 path (three in-place sweeps) on the same grid
    K3 = min(K, a.three_sweep_steps)
    g64.set_time_blocking(0)
    g64.zero_fields()
    r3 = run_path(g64, K3, 2, path=_lib.PATH_3SWEEP)
    log(f"fp64 3-sweep: {r3['gcells']:.1f} Gcell/s")
    g64.close()


    def kernel_names(prec, cfg=None):
        sname = {Precision.FP32: "float", Precision.FP64: "double", Precision.FP16: "_Float16"}[prec]
        cname = "double" if prec == Precision.FP64 else "float"
        if cfg is None:
            return f"fdtd::fused_sweep<{sname},{cname},{sname},V,true>"
        if cfg.get("R", 0) > 0:
            return f"fdtd::strip_sweep<{sname},{cname},{cfg['V']},{cfg['R']},{cfg['L']},false>"
        return f"fdtd::blocked_sweep<{sname},{cname},{cfg['V']},{cfg['L']},false>"

    def roofline(prec, blocked):
        """Roofline of the steady kernel of a path, from its launches inside
        the timed region (HIP events on the grid's stream).  Algorithmic
        bytes per launch: E and B read once and written once, 12 words per
        cell (a temporally blocked pass advances L steps over that one
        read and write).  `traffic`: HBM bytes per launch measured with
        rocprofv3 counters for the same kernel and geometry (committed under
        profiles/), `traffic_frac` the same over the launch time.  For
        blocked passes `effective_frac` restates the launch as L unblocked
        steps (12 words per cell-update each), a figure that can exceed 1."""
        cfg = blocking[prec] if blocked else None
        r = best[prec] if blocked else fused[prec]
        word = WORD[prec]
        alg = 12.0 * word * cells
        ms = r["steady_launch_ms"]
        gbs = alg / (ms * 1e-3) / 1e9
        prof = committed_profile(prec, n, cfg, tunes[prec])
        out = {"bound": "hbm", "achieved": gbs, "peak": PEAK_GBS, "unit": "GB/s",
               "frac": gbs / PEAK_GBS,
               "traffic": prof.get("hbm_bytes_per_launch") if prof else None,
               "kernel": kernel_names(prec, cfg), "alg_bytes_per_launch": alg, "launch_ms": ms,
               "launches_timed": (r["kernels"]["steady"][1] if r.get("kernels") else None)}
        if prof:
            out["traffic_frac"] = prof["hbm_bytes_per_launch"] / (ms * 1e-3) / 1e9 / PEAK_GBS
            out["traffic_profile"] = prof.get("file")
            if prof.get("valu_busy") is not None:
                out["valu_busy"] = prof["valu_busy"]
        if blocked:
            L = cfg["L"]
            out["steps_per_launch"] = L
            out["effective_achieved"] = gbs * L
            out["effective_frac"] = gbs * L / PEAK_GBS
            if r.get("last_launch_ms"):
                out["last_launch_ms"] = r["last_launch_ms"]
        return out

    def path_rec(r, prec, tune=None, bytes_per_cell=None, cfg=None):
        bpc = bytes_per_cell or 12 * WORD[prec]
        gbs = bpc * cells / (r["ms_per_step"] * 1e-3) / 1e9
        rec = dict(gcells=r["gcells"], ms_per_step=r["ms_per_step"], launch_ms=r["launch_ms"],
                   alg_bytes_per_cell=bpc, alg_gbs_per_step=gbs, frac_of_peak=gbs / PEAK_GBS)
        if tune:
            rec["launch_shape"] = dict(V=tune[0], TY=tune[1], kchunk=tune[2])
        if cfg:
            rec["blocking"] = cfg
        return rec

    blocked32 = blocking[Precision.FP32]["L"] > 0
    r32 = best[Precision.FP32]
    cpu = None if a.no_cpu_baseline else cpu_baseline(n)
    cfg32 = blocking[Precision.FP32]
    out = {
        "metric": METRIC,
        "value": r32["gcells"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{n}^3 periodic vacuum Yee grid, dx=1, dt=0.5dx/(c*sqrt3), "
                         "Gaussian-modulated sinusoid Jz point source at the centre "
                         "(period 20dt, sigma 40dt, peak 120dt), zero initial fields, "
                         f"{W} untimed steps, then {K} timed steps continuing the run"),
            "grid": [n, n, n],
            "path": ("fp32 storage and compute, temporally blocked passes of "
                     f"{cfg32['L']} fused steps (autotuned; bit-identical to single sweeps)"
                     if blocked32 else "fp32 storage and compute, fused E-then-double-half-B sweep"),
            "blocking": cfg32 if blocked32 else None,
            "fused_launch_shape": dict(V=tunes[Precision.FP32][0], TY=tunes[Precision.FP32][1],
                                       kchunk=tunes[Precision.FP32][2]),
        },
        "roofline": roofline(Precision.FP32, blocked32),
        "fused_sweep_roofline": roofline(Precision.FP32, False),
        "rooflines": {"fp64": roofline(Precision.FP64, blocking[Precision.FP64]["L"] > 0),
                      "fp16": roofline(Precision.FP16, blocking[Precision.FP16]["L"] > 0),
                      "fp64_fused": roofline(Precision.FP64, False),
                      "fp16_fused": roofline(Precision.FP16, False)},
        "cpu_baseline": cpu,
        "paths": {
            "fp32": path_rec(best[Precision.FP32], Precision.FP32, cfg=blocking[Precision.FP32]),
            "fp64": path_rec(best[Precision.FP64], Precision.FP64, cfg=blocking[Precision.FP64]),
            "fp16": path_rec(best[Precision.FP16], Precision.FP16, cfg=blocking[Precision.FP16]),
            "fp32_fused": path_rec(fused[Precision.FP32], Precision.FP32, tunes[Precision.FP32]),
            "fp64_fused": path_rec(fused[Precision.FP64], Precision.FP64, tunes[Precision.FP64]),
            "fp16_fused": path_rec(fused[Precision.FP16], Precision.FP16, tunes[Precision.FP16]),
            "fp64_3sweep": path_rec(r3, Precision.FP64, bytes_per_cell=240),
        },
        "speedup_fp32_vs_fp64": r32["gcells"] / best[Precision.FP64]["gcells"],
        "speedup_fp32_fused_vs_fp64_fused": (fused[Precision.FP32]["gcells"]
                                             / fused[Precision.FP64]["gcells"]),
        "speedup_fp32_vs_fp64_3sweep": r32["gcells"] / r3["gcells"],
        "speedup_blocked_vs_fused_fp32": r32["gcells"] / fused[Precision.FP32]["gcells"],
        "errors_vs_fp64": {"fp32": e32, "fp16": e16},
        "energy": en,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""The seeded synthetic workloads the benchmarks and the full-size tests run.

Every workload is a periodic vacuum grid with dx = dy = dz = 1 and
dt = 0.5 dx / (c sqrt 3) (Courant factor 0.5), driven by point current
sources sampled at t = (n+1) dt before step n, as the reference driver
samples its current (perf-tests/sample/sample.cpp:74):

* ``centre_source``: the 512^3 headline source, a Gaussian-modulated
  sinusoid on Jz at the grid centre (period 20 dt, envelope sigma 40 dt,
  peak at 120 dt, amplitude 1);
* ``precision_waves``: 8 seeded plane waves (integer wave numbers in
  [-8, 8], amplitude U[0.1, 1], phase U[0, 2 pi), random polarization) that
  the 768^3 precision sweep superposes on the centre source;
* ``anisotropic_sources``: 4 seeded Jz point sources with distinct integer
  periods in [16, 40] dt and no envelope, for the 1024x1024x256 grid.
"""
from __future__ import annotations

import numpy as np

from .structures import C

DT = 0.5 / (C * np.sqrt(3.0))


def centre_source(n, dt=DT, amp=1.0):
    """Gaussian-modulated Jz point source at (n/2, n/2, n/2)."""
    return dict(i=n // 2, j=n // 2, k=n // 2, comp=2, kind=1, amp=amp, period=20 * dt,
                t0=120 * dt, sigma=40 * dt, phase=0.0)


def precision_waves(seed=2024, count=8):
    """The precision sweep's seeded plane waves (arguments of add_plane_wave)."""
    rng = np.random.default_rng(seed)
    waves = []
    while len(waves) < count:
        m = [int(x) for x in rng.integers(-8, 9, 3)]
        if not any(m):
            continue
        waves.append(dict(m=m, amp=float(rng.uniform(0.1, 1.0)),
                          phase=float(rng.uniform(0, 2 * np.pi)),
                          pol=[float(x) for x in rng.standard_normal(3)]))
    return waves


def anisotropic_sources(dims, seed=2024, dt=DT):
    """(sources, periods in dt): 4 Jz point sources at seeded cells."""
    Ni, Nj, Nk = dims
    rng = np.random.default_rng(seed)
    periods = rng.permutation(np.arange(16, 41))[:4]
    srcs = [dict(i=int(rng.integers(Ni)), j=int(rng.integers(Nj)), k=int(rng.integers(Nk)),
                 comp=2, kind=1, amp=1.0, period=float(p) * dt, t0=0.0, sigma=0.0)
            for p in periods]
    return srcs, [int(p) for p in periods]


def add_waves(g, waves):
    for w in waves:
        g.add_plane_wave(w["m"], w["amp"], w["phase"], w["pol"])


def error_summary(test, ref):
    """Relative L2 (all six components together) and per-component
    max-abs error of grid `test` against the FP64 grid `ref`, plus each
    component's max |ref| estimate (sqrt of its mean square)."""
    per = test.compare(ref)
    names = ["Ex", "Ey", "Ez", "Bx", "By", "Bz"]
    sr_all = sum(p[1] for p in per)
    comp = {}
    for nm, (sd, sr, mx) in zip(names, per):
        comp[nm] = dict(rel_l2_of_total=float(np.sqrt(sd / sr_all)) if sr_all > 0 else None,
                        max_abs=float(mx), rms_ref=float(np.sqrt(sr / test.cells)))
    tot = float(np.sqrt(sum(p[0] for p in per) / sr_all)) if sr_all > 0 else 0.0
    return dict(rel_l2=tot, max_abs=max(p[2] for p in per),
                rms_ref=max(float(np.sqrt(p[1] / test.cells)) for p in per), per_component=comp)


def energy(g):
    """Field energy sum (E^2 + B^2) / (8 pi) over the grid (per unit cell volume)."""
    return float(sum(g.sum_squares())) / (8.0 * np.pi)

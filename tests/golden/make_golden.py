"""Write the golden fixtures from the reference's own compiled class.

Run in a checkout where oracle/_ref/libfdtd_ref.so was built from the
reference sources (``make -C oracle ref``):

    python tests/golden/make_golden.py

Fixtures (data only: inputs and the reference's outputs):

* ref_16x12x10_random_50steps.npz -- seeded random E, B and a persistent
  current Jx = Jy = Jz on five cells (the J pattern on which the reference's
  OpenMP E update, src/FDTD/FDTD.cpp:85-93, and its Kokkos functor,
  kokkos_functors.h:81-89, agree), 50 update_fields() calls, full dump.
* ref_config1_64cube_200steps.json -- 64^3 x-propagating plane wave
  (Ey = sin(2 pi x/64) at x = i, Bz = Ey/c at x = i + 1/2, dx = 1,
  dt = 0.5 dx/(c sqrt 3)), 200 steps: SHA-256 and L2 norm per component.
* ref_sample_32_100.json -- the reference driver's spherical_wave(32, 100)
  (perf-tests/sample/sample.cpp:14-91) driven through the class exactly as
  the driver does: SHA-256 per component and the printed 10x10 Ex slice.
* ref_pml_16x12x10_random_40steps.npz -- the PML class
  (src/FDTD/FDTD_PML.cpp, run on one thread) with pml_percent 0.25 on
  seeded random E, B and a persistent Jx = Jy = Jz on cells of the main box
  and of the layer, 40 update_fields() calls, full dump of the totals.
* ref_pml_sample_32_100.json -- the driver's PML block
  (perf-tests/sample/sample.cpp:93-123, FDTD_PML with pml_percent 0.2) for
  spherical_wave(32, 100): SHA-256 per component and the 10x10 Ex slice
  (:136-147).

``python tests/golden/make_golden.py [case ...]`` writes the named cases
(random, config1, sample, pml_random, pml_sample; default all).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def random_case():
    Ni, Nj, Nk, nsteps = 16, 12, 10, 50
    dx, dy, dz = 1.0, 1.25, 0.8
    dt = 0.45 * dz / (O.C * np.sqrt(3.0))
    n = Ni * Nj * Nk
    rng = np.random.default_rng(20241202)
    E = [rng.standard_normal(n) for _ in range(3)]
    B = [rng.standard_normal(n) for _ in range(3)]
    J = np.zeros(n)
    cells = np.array([0, 37, 500, 1234, n - 1])
    J[cells] = rng.standard_normal(cells.size)
    g = O.RefGrid(Ni, Nj, Nk, dx, dy, dz, dt)
    for c in range(3):
        g.set(c, E[c])
        g.set(3 + c, B[c])
        g.set(6 + c, J)
    g.update_fields(nsteps)
    out = [g.get(c) for c in range(6)]
    np.savez_compressed(
        os.path.join(HERE, "ref_16x12x10_random_50steps.npz"),
        dims=np.array([Ni, Nj, Nk]), spacing=np.array([dx, dy, dz]), dt=np.array(dt),
        nsteps=np.array(nsteps), E_in=np.array(E), B_in=np.array(B), J_in=J,
        E_out=np.array(out[:3]), B_out=np.array(out[3:]))


def config1_state(N=64):
    """Initial fields of the 64^3 plane-wave case (x = i for Ey, i + 1/2 for Bz)."""
    n = N ** 3
    i = np.arange(n) % N
    ey = np.sin(2.0 * np.pi * i / N)
    bz = np.sin(2.0 * np.pi * (i + 0.5) / N) / O.C
    z = np.zeros(n)
    return [z.copy(), ey, z.copy(), z.copy(), z.copy(), bz]


def config1_case():
    N, nsteps = 64, 200
    d = 1.0
    dt = 0.5 * d / (O.C * np.sqrt(3.0))
    state = config1_state(N)
    g = O.RefGrid(N, N, N, d, d, d, dt)
    for c in range(6):
        g.set(c, state[c])
    g.update_fields(nsteps)
    out = [g.get(c) for c in range(6)]
    rec = dict(dims=[N, N, N], d=d, dt=dt, nsteps=nsteps,
               init="Ey=sin(2*pi*i/N), Bz=sin(2*pi*(i+0.5)/N)/C (numpy sin), others 0",
               init_sha256=[sha(a) for a in state],
               sha256=[sha(a) for a in out],
               l2=[float(np.sqrt(np.sum(a * a))) for a in out])
    with open(os.path.join(HERE, "ref_config1_64cube_200steps.json"), "w") as f:
        json.dump(rec, f, indent=1)


def sample_case(n=32, it=100):
    p = O.sample_box(n)
    g = O.RefGrid(n, n, n, p["d"], p["d"], p["d"], p["dt"], -p["boundary"], p["boundary"],
                  -p["boundary"], p["boundary"], -p["boundary"], p["boundary"])
    cur_time = min(p["iterations"], it)
    for t in range(cur_time):
        for k in range(p["start"], p["stop"]):
            for j in range(p["start"], p["stop"]):
                for i in range(p["start"], p["stop"]):
                    idx = i + j * n + k * n * n
                    val = O.sample_current(float(i) * p["d"], float(j) * p["d"],
                                           float(k) * p["d"], float(t + 1) * p["dt"],
                                           p["T"], p["Tx"], p["Ty"], p["Tz"])
                    for comp in (6, 7, 8):
                        g.set_value(comp, idx, val)
        g.update_fields(1)
    g.zeroed_currents()
    g.update_fields(it - cur_time)
    out = [g.get(c) for c in range(6)]
    k = n // 2
    sl = [[float(out[0][i + j * n + k * n * n]) for i in range(n // 2 - 5, n // 2 + 5)]
          for j in range(n // 2 - 5, n // 2 + 5)]
    rec = dict(n=n, iterations=it, sha256=[sha(a) for a in out],
               l2=[float(np.sqrt(np.sum(a * a))) for a in out], ex_slice=sl)
    with open(os.path.join(HERE, f"ref_sample_{n}_{it}.json"), "w") as f:
        json.dump(rec, f, indent=1)


def pml_random_case():
    Ni, Nj, Nk, nsteps, pml = 16, 12, 10, 40, 0.25
    dx, dy, dz = 1.0, 1.25, 0.8
    dt = 0.45 * dz / (O.C * np.sqrt(3.0))
    n = Ni * Nj * Nk
    rng = np.random.default_rng(20241203)
    E = [rng.standard_normal(n) for _ in range(3)]
    B = [rng.standard_normal(n) for _ in range(3)]
    J = np.zeros(n)
    # (8,6,5) and (9,5,4) lie in the main box, the others in the layer
    cells = np.array([8 + 6 * Ni + 5 * Ni * Nj, 9 + 5 * Ni + 4 * Ni * Nj, 0, 1 + 2 * Ni, n - 1])
    J[cells] = rng.standard_normal(cells.size)
    g = O.RefPmlGrid(Ni, Nj, Nk, dx, dy, dz, dt, pml)
    for c in range(3):
        g.set(c, E[c])
        g.set(3 + c, B[c])
        g.set(6 + c, J)
    g.update_fields(nsteps)
    out = [g.get(c) for c in range(6)]
    np.savez_compressed(
        os.path.join(HERE, "ref_pml_16x12x10_random_40steps.npz"),
        dims=np.array([Ni, Nj, Nk]), spacing=np.array([dx, dy, dz]), dt=np.array(dt),
        pml_percent=np.array(pml), nsteps=np.array(nsteps), E_in=np.array(E), B_in=np.array(B),
        J_in=J, E_out=np.array(out[:3]), B_out=np.array(out[3:]))


def pml_sample_case(n=32, it=100, pml=0.2):
    p = O.sample_box(n)
    g = O.RefPmlGrid(n, n, n, p["d"], p["d"], p["d"], p["dt"], pml, -p["boundary"],
                     p["boundary"], -p["boundary"], p["boundary"], -p["boundary"], p["boundary"])
    cur_time = min(p["iterations"], it)
    for t in range(cur_time):
        for k in range(p["start"], p["stop"]):
            for j in range(p["start"], p["stop"]):
                for i in range(p["start"], p["stop"]):
                    idx = i + j * n + k * n * n
                    val = O.sample_current(float(i) * p["d"], float(j) * p["d"],
                                           float(k) * p["d"], float(t + 1) * p["dt"],
                                           p["T"], p["Tx"], p["Ty"], p["Tz"])
                    for comp in (6, 7, 8):
                        g.set_value(comp, idx, val)
        g.update_fields(1)
    g.zeroed_currents()
    g.update_fields(it - cur_time)
    out = [g.get(c) for c in range(6)]
    k = n // 2
    sl = [[float(out[0][i + j * n + k * n * n]) for i in range(n // 2 - 5, n // 2 + 5)]
          for j in range(n // 2 - 5, n // 2 + 5)]
    rec = dict(n=n, iterations=it, pml_percent=pml, sha256=[sha(a) for a in out],
               l2=[float(np.sqrt(np.sum(a * a))) for a in out], ex_slice=sl)
    with open(os.path.join(HERE, f"ref_pml_sample_{n}_{it}.json"), "w") as f:
        json.dump(rec, f, indent=1)


CASES = dict(random=random_case, config1=config1_case, sample=sample_case,
             pml_random=pml_random_case, pml_sample=pml_sample_case)

if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libfdtd_ref.so not built (make -C oracle ref)")
    for name in (sys.argv[1:] or list(CASES)):
        CASES[name]()
    print("golden fixtures written to", HERE)

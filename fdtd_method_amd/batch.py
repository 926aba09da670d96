"""Batched solves: many independent grids advanced by one launch per step.

``Batch(grids)`` binds the C-ABI batch (include/fdtd_mi355x.h,
fdtd_batch_*): each step of every grid is the reference's update_fields()
(src/FDTD/FDTD.cpp:153-157); one persistent launch per step walks a work
queue over the tiles of all grids.  ``plane_wave_batch`` builds the seeded
many-grid workload (uneven dimensions, one plane wave and one
Gaussian-modulated Jz source per grid).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .fdtd import FDTD
from .structures import C, Parameters, Precision


class Batch:
    def __init__(self, grids):
        self.grids = list(grids)
        L = _lib.load()
        arr = (ctypes.c_void_p * len(self.grids))(*[g.handle.value for g in self.grids])
        h = ctypes.c_void_p()
        _lib.check(L.fdtd_batch_create(len(self.grids), arr, ctypes.byref(h)))
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.fdtd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_grids(self):
        for g in self.grids:
            g._flush()

    def _after(self):
        for g in self.grids:
            g._invalidate()

    def update_fields(self, nsteps: int) -> None:
        self._sync_grids()
        _lib.check(self._L.fdtd_batch_update_fields(self._h, int(nsteps)))
        self._after()

    def update_fields_graph(self, nsteps: int, replays: int = 1) -> None:
        """nsteps*replays steps: capture nsteps once as a HIP graph, replay it."""
        self._sync_grids()
        _lib.check(self._L.fdtd_batch_update_fields_graph(self._h, int(nsteps), int(replays)))
        self._after()

    def time(self, nsteps: int, sequential: bool = True):
        """(ms per step batched, ms per step with one launch per grid)."""
        self._sync_grids()
        a, b = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._L.fdtd_batch_time(self._h, int(nsteps), ctypes.byref(a),
                                           ctypes.byref(b) if sequential else None))
        self._after()
        return a.value, (b.value if sequential else None)

    @property
    def shape(self):
        out = (ctypes.c_longlong * 4)()
        _lib.check(self._L.fdtd_batch_get_shape(self._h, out))
        return dict(V=out[0], threads=out[1], tile_cells=out[2], work_items=out[3])

    def set_shape(self, V: int, threads: int, tile_cells: int) -> None:
        _lib.check(self._L.fdtd_batch_set_shape(self._h, int(V), int(threads), int(tile_cells)))

    def autotune(self, reps: int = 3):
        out = (ctypes.c_longlong * 4)()
        ms = ctypes.c_float()
        _lib.check(self._L.fdtd_batch_autotune(self._h, int(reps), out, ctypes.byref(ms)))
        return dict(V=out[0], threads=out[1], tile_cells=out[2], work_items=out[3]), ms.value

    def set_time_blocking_strips(self, L: int, V: int = 0, R: int = 0, W: int = 0) -> None:
        """Runs of at least L steps advance in persistent strip passes of L
        steps (0 picks V=1, R=2, W=16; L=0 turns it off); bit-identical to
        single-grid runs."""
        _lib.check(self._L.fdtd_batch_set_time_blocking_strips(self._h, int(L), int(V), int(R),
                                                               int(W)))

    def get_time_blocking_strips(self):
        out = (ctypes.c_int * 5)()
        _lib.check(self._L.fdtd_batch_get_time_blocking_strips(self._h, out))
        return dict(L=out[0], V=out[1], R=out[2], W=out[3], work_items=out[4])

    @property
    def cells(self) -> int:
        return sum(g.cells for g in self.grids)


def yee_plane_wave(shape, m, amp, phase, pol):
    """Host (numpy) form of FDTD.add_plane_wave: E = amp*e*cos(k.r+phase),
    B = khat x E, with integer wave numbers m (periodic on the grid), each
    component at its Yee position (Ex at (i+1/2,j,k), Ey at (i,j+1/2,k), Ez
    at (i,j,k+1/2), Bx at (i,j+1/2,k+1/2), By at (i+1/2,j,k+1/2), Bz at
    (i+1/2,j+1/2,k)); dx = 1, Gaussian units."""
    Ni, Nj, Nk = shape
    k = 2.0 * np.pi * np.array([m[0] / Ni, m[1] / Nj, m[2] / Nk])
    khat = k / np.linalg.norm(k)
    v = np.asarray(pol, dtype=float)
    e = v - np.dot(v, khat) * khat
    e /= np.linalg.norm(e)
    b = np.cross(khat, e)
    k_i, j_i, i_i = np.meshgrid(np.arange(Nk), np.arange(Nj), np.arange(Ni), indexing="ij")
    offs = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)]
    out = []
    for c, (ox, oy, oz) in enumerate(offs):
        ph = k[0] * (i_i + ox) + k[1] * (j_i + oy) + k[2] * (k_i + oz) + phase
        pol = e[c] if c < 3 else b[c - 3]
        out.append((amp * pol * np.cos(ph)).reshape(-1))
    return out


def plane_wave_batch(n_grids=256, seed=2024, dims=(32, 160, 16), precision=Precision.FP32,
                     lam=(8.0, 32.0), with_sources=True):
    """The seeded many-grid workload: each dimension uniform in
    {dims[0], dims[0]+dims[2], ..., dims[1]}, dx = 1, dt = 0.5/(c sqrt 3),
    one plane wave (random integer wave vector with wavelength in lam cells,
    amplitude U[0.1, 1], phase U[0, 2 pi), random polarization), generated
    on the device, and one Gaussian-modulated Jz
    point source at a random cell (period U[16, 40] dt, sigma 40 dt, peak
    120 dt).  Returns (grids, description)."""
    rng = np.random.default_rng(seed)
    choices = np.arange(dims[0], dims[1] + 1, dims[2])
    dt = 0.5 / (C * np.sqrt(3.0))
    grids, desc = [], []
    for q in range(n_grids):
        shape = tuple(int(x) for x in rng.choice(choices, 3))
        while True:
            m = rng.integers(-4, 5, 3)
            if not m.any():
                continue
            kk = np.sqrt(sum((m[a] / shape[a]) ** 2 for a in range(3)))
            if lam[0] <= 1.0 / kk <= lam[1]:
                break
        amp = float(rng.uniform(0.1, 1.0))
        phase = float(rng.uniform(0.0, 2.0 * np.pi))
        g = FDTD(Parameters(*shape), dt, precision)
        pol = np.random.default_rng(seed * 1000 + q).standard_normal(3)
        g.add_plane_wave(m, amp, phase, pol)
        src = None
        if with_sources:
            src = dict(i=int(rng.integers(shape[0])), j=int(rng.integers(shape[1])),
                       k=int(rng.integers(shape[2])), comp=2, kind=1, amp=1.0,
                       period=float(rng.uniform(16, 40)) * dt, t0=120 * dt, sigma=40 * dt)
            g.set_point_sources([src])
        grids.append(g)
        desc.append(dict(shape=shape, m=[int(x) for x in m], amp=amp, phase=phase, source=src))
    return grids, desc

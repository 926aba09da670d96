"""Python mirror of the reference's FDTD class on the MI355X library.

``FDTD`` has the public surface of ``FDTD_openmp::FDTD`` /
``FDTD_kokkos::FDTD`` (include/FDTD/FDTD.h:8-41,
include/FDTD_kokkos/FDTD_kokkos.h:7-36):

=====================================  ==========================================
reference                              here
=====================================  ==========================================
``FDTD(Parameters, FP dt)``            ``FDTD(parameters, dt, precision=FP64)``
``Field& get_field(Component)``        ``get_field(component) -> Field``
``virtual void update_fields()``       ``update_fields(nsteps=1)``
``void zeroed_currents()``             ``zeroed_currents()``
``FDTD_PML(Parameters, FP dt, FP pml)`` ``FDTD_PML(parameters, dt, pml_percent, precision=FP64)``
=====================================  ==========================================

Errors match the reference's: ``ValueError("ERROR: invalid parameters")``
(std::invalid_argument, src/FDTD/FDTD.cpp:6) and
``LookupError("ERROR: Invalid field component")`` (std::logic_error, :149).

``Field`` is a host mirror of one device array supporting ``f[idx]``
(OpenMP style) and ``f(idx)`` (Kokkos style).  Reads download lazily;
writes to E/B mark the mirror dirty and are uploaded before the next step;
single-cell writes to J go to the device as sparse current values (the
drivers' 8-cell box source, perf-tests/sample/sample.cpp:67-81).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .structures import Component, Parameters, Precision


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Field:
    """Host mirror of one component (the reference's ``Field&``)."""

    def __init__(self, owner: "FDTD", comp: Component):
        self._owner = owner
        self._comp = Component(comp)
        self._host: np.ndarray | None = None
        self._dirty = False

    # -- mirror management ---------------------------------------------------
    def _pull(self) -> np.ndarray:
        if self._host is None:
            self._host = self._owner.download(self._comp)
        return self._host

    def _invalidate(self) -> None:
        if not self._dirty:
            self._host = None

    def _push(self) -> None:
        if self._dirty:
            self._owner.upload(self._comp, self._host)
            self._dirty = False

    # -- element access ------------------------------------------------------
    def __len__(self) -> int:
        return self._owner.cells

    def size(self) -> int:
        return self._owner.cells

    def __getitem__(self, idx):
        return self._pull()[idx]

    def __call__(self, idx):
        return self._pull()[idx]

    def __setitem__(self, idx, value):
        if self._comp >= Component.JX and np.isscalar(idx) and np.isscalar(value) \
                and not self._dirty:
            # sparse current write: goes straight to the device list
            self._owner.set_current_values([int(idx)], [int(self._comp)], [float(value)])
            if self._host is not None:
                self._host[idx] = value
            return
        host = self._pull()
        host[idx] = value
        self._dirty = True

    def numpy(self) -> np.ndarray:
        """Copy of the current values (FP64)."""
        self._owner._flush()
        return self._pull().copy()

    def assign(self, values) -> None:
        """Replace the whole field."""
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != self._owner.cells:
            raise ValueError("size mismatch")
        self._host = a.copy()
        self._dirty = True


class DeviceArray:
    """One device array of a grid: address ``ptr``, element type ``dtype``,
    ``size`` elements (index i + j*Ni + k*Ni*Nj) and FP16 storage ``scale``.
    Exposes ``__cuda_array_interface__``, so ``torch.as_tensor(a,
    device="cuda")`` aliases it; it keeps its grid alive."""

    def __init__(self, ptr: int, dtype: np.dtype, size: int, scale: float, owner: "FDTD"):
        self.ptr, self.dtype, self.size, self.scale = ptr, dtype, size, scale
        self._owner = owner

    @property
    def __cuda_array_interface__(self):
        return {"shape": (self.size,), "typestr": self.dtype.str, "data": (self.ptr, False),
                "version": 2}


class FDTD:
    """Yee solver on one MI355X grid (reference API, see module docstring)."""

    def __init__(self, parameters: Parameters, dt: float, precision: Precision = Precision.FP64):
        self._open(parameters, dt, precision, None)

    def _open(self, parameters, dt, precision, pml_percent) -> None:
        L = _lib.load()
        p = _lib.Parameters(parameters.Ni, parameters.Nj, parameters.Nk, parameters.ax,
                            parameters.bx, parameters.ay, parameters.by, parameters.az,
                            parameters.bz, parameters.dx, parameters.dy, parameters.dz)
        h = ctypes.c_void_p()
        if pml_percent is None:
            _lib.check(L.fdtd_create(ctypes.byref(p), float(dt), int(precision), ctypes.byref(h)))
        else:
            _lib.check(L.fdtd_create_pml(ctypes.byref(p), float(dt), float(pml_percent),
                                         int(precision), ctypes.byref(h)))
        self._h = h
        self._L = L
        self.parameters = parameters
        self.dt = float(dt)
        self.precision = Precision(precision)
        self.shape = (parameters.Ni, parameters.Nj, parameters.Nk)
        self.cells = parameters.Ni * parameters.Nj * parameters.Nk
        self._fields = {}
        self._device = False  # device addresses handed out: mirrors refresh after every call
        # the update_E/update_B box, include/FDTD/FDTD.h:17-18 (whole grid here;
        # FDTD_PML narrows it to the interior as src/FDTD/FDTD_PML.cpp:263-268)
        self.begin_main = (0, 0, 0)
        self.end_main = self.shape

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.fdtd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- reference API ---------------------------------------------------------
    def get_field(self, component) -> Field:
        try:
            comp = Component(component)
        except ValueError:
            raise LookupError("ERROR: Invalid field component") from None
        f = self._fields.get(comp)
        if f is None:
            f = self._fields[comp] = Field(self, comp)
        return f

    def update_fields(self, nsteps: int = 1) -> None:
        """nsteps reference update_fields() calls (fused sweeps)."""
        self._flush()
        _lib.check(self._L.fdtd_update_fields(self._h, int(nsteps)))
        self._invalidate()

    def update_fields_3sweep(self, nsteps: int = 1) -> None:
        """Same steps in the reference's B/2, E, B/2 three-sweep structure."""
        self._flush()
        _lib.check(self._L.fdtd_update_fields_3sweep(self._h, int(nsteps)))
        self._invalidate()

    def zeroed_currents(self) -> None:
        for c in (Component.JX, Component.JY, Component.JZ):
            f = self._fields.get(c)
            if f is not None:
                f._host = None
                f._dirty = False
        _lib.check(self._L.fdtd_zero_currents(self._h))

    # -- extensions ---------------------------------------------------------------
    def _flush(self) -> None:
        for f in self._fields.values():
            f._push()

    def _invalidate(self) -> None:
        for c, f in self._fields.items():
            if c < Component.JX or self._device:
                f._invalidate()

    def upload(self, comp, values) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != self.cells:
            raise ValueError(f"expected {self.cells} values, got {a.size}")
        _lib.check(self._L.fdtd_upload(self._h, int(comp), _dptr(a), a.size))

    def download(self, comp) -> np.ndarray:
        out = np.empty(self.cells, dtype=np.float64)
        _lib.check(self._L.fdtd_download(self._h, int(comp), _dptr(out), out.size))
        return out

    def download_plane(self, comp, k: int) -> np.ndarray:
        """FP64 copy of plane k of a component, shape (Nj, Ni)."""
        self._flush()
        Ni, Nj, _ = self.shape
        out = np.empty(Ni * Nj, dtype=np.float64)
        _lib.check(self._L.fdtd_download_plane(self._h, int(comp), int(k), _dptr(out), out.size))
        return out.reshape(Nj, Ni)

    def set_current_values(self, index, comp, value) -> None:
        idx = np.ascontiguousarray(index, dtype=np.int64)
        cc = np.ascontiguousarray(comp, dtype=np.int32)
        vv = np.ascontiguousarray(value, dtype=np.float64)
        _lib.check(self._L.fdtd_set_current_values(
            self._h, idx.size, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            cc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(vv)))

    def update_fields_with_currents(self, writes) -> None:
        """len(writes) steps; before step s the sparse currents writes[s] =
        (index, comp, value) (sequences of equal length, comp 6..8) are set as
        set_current_values would: the reference drivers' loop of J writes and
        update_fields() as one call (fdtd_update_fields_with_currents)."""
        self._flush()
        counts = np.array([len(w[0]) for w in writes], dtype=np.int32)
        idx = np.concatenate([np.asarray(w[0], dtype=np.int64) for w in writes] or
                             [np.zeros(0, np.int64)])
        cc = np.concatenate([np.asarray(w[1], dtype=np.int32) for w in writes] or
                            [np.zeros(0, np.int32)])
        vv = np.concatenate([np.asarray(w[2], dtype=np.float64) for w in writes] or
                            [np.zeros(0)])
        idx, cc, vv = (np.ascontiguousarray(a) for a in (idx, cc, vv))
        _lib.check(self._L.fdtd_update_fields_with_currents(
            self._h, len(writes), counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            cc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dptr(vv)))
        self._invalidate()
        for c in (Component.JX, Component.JY, Component.JZ):
            f = self._fields.get(c)
            if f is not None and not f._dirty:
                f._host = None

    def launch_stats(self):
        """{steps, passes, sweeps, prologues} launched since creation."""
        out = (ctypes.c_longlong * 4)()
        _lib.check(self._L.fdtd_get_stats(self._h, out))
        return dict(steps=out[0], passes=out[1], sweeps=out[2], prologues=out[3])

    def set_probes(self, cells, capacity: int) -> None:
        """Record FP64 values of single cells after every step on the device.

        cells: iterable of (i, j, k, component) with component 0..5 (Ex..Bz);
        capacity: rows (steps) the device buffer holds.  An empty list removes
        the probes."""
        cells = list(cells)
        n = len(cells)
        idx = (ctypes.c_longlong * max(1, n))()
        comp = (ctypes.c_int * max(1, n))()
        for q, (i, j, k, c) in enumerate(cells):
            Ni, Nj, _ = self.shape
            idx[q] = int(i) + int(j) * Ni + int(k) * Ni * Nj
            comp[q] = int(c)
        _lib.check(self._L.fdtd_set_probes(self._h, n, idx, comp, int(capacity)))
        self._nprobe = n

    def probe_rows(self) -> int:
        out = ctypes.c_longlong()
        _lib.check(self._L.fdtd_probe_rows(self._h, ctypes.byref(out)))
        return out.value

    def read_probes(self, row0: int = 0, nrows=None) -> np.ndarray:
        """Probe rows [row0, row0 + nrows) as an (nrows, nprobes) FP64 array."""
        if nrows is None:
            nrows = self.probe_rows() - row0
        out = np.empty((nrows, getattr(self, "_nprobe", 0)), dtype=np.float64)
        _lib.check(self._L.fdtd_read_probes(self._h, int(row0), int(nrows), _dptr(out)))
        return out

    def set_energy_interval(self, every: int, capacity: int = 0) -> None:
        """Sums of squares of Ex..Bz after every step that is a multiple of
        `every`, reduced on the device into a buffer of `capacity` rows."""
        _lib.check(self._L.fdtd_set_energy_interval(self._h, int(every), int(capacity)))

    def energy_records(self):
        """(steps, sums): the recorded steps and an (n, 6) array of sums."""
        n = ctypes.c_longlong()
        _lib.check(self._L.fdtd_energy_rows(self._h, ctypes.byref(n)))
        steps = (ctypes.c_longlong * max(1, n.value))()
        sums = np.empty((n.value, 6), dtype=np.float64)
        _lib.check(self._L.fdtd_read_energy(self._h, 0, n.value, steps, _dptr(sums)))
        return [steps[r] for r in range(n.value)], sums

    def rewind_records(self) -> None:
        """Probe rows and energy records restart at row 0."""
        _lib.check(self._L.fdtd_rewind_records(self._h))

    def count_nonfinite(self):
        """Number of NaN/Inf values of Ex..Bz as stored (list of 6)."""
        self._flush()
        out = (ctypes.c_longlong * 6)()
        _lib.check(self._L.fdtd_count_nonfinite(self._h, out))
        return list(out)

    def set_point_sources(self, sources) -> None:
        """sources: iterable of dicts {i,j,k,comp(0..2),kind,amp,period,t0,sigma,phase}."""
        arr = (_lib.PointSource * max(1, len(sources)))()
        for q, s in enumerate(sources):
            arr[q] = _lib.PointSource(int(s["i"]), int(s["j"]), int(s["k"]), int(s["comp"]),
                                      int(s.get("kind", 1)), 0, float(s.get("amp", 1.0)),
                                      float(s.get("period", 1.0)), float(s.get("t0", 0.0)),
                                      float(s.get("sigma", 0.0)), float(s.get("phase", 0.0)))
        _lib.check(self._L.fdtd_set_point_sources(self._h, len(sources), arr))

    @property
    def step(self) -> int:
        v = ctypes.c_longlong()
        _lib.check(self._L.fdtd_get_step(self._h, ctypes.byref(v)))
        return v.value

    @step.setter
    def step(self, n: int) -> None:
        _lib.check(self._L.fdtd_set_step(self._h, int(n)))

    def synchronize(self) -> None:
        _lib.check(self._L.fdtd_synchronize(self._h))
        if self._device:
            self._invalidate()

    # -- device access and the protected curl updates --------------------------
    def device_field(self, component) -> "DeviceArray":
        """The device array of a component (0..8), for the caller's own GPU
        work: the Kokkos class's ``get_field`` returns device Views
        (include/FDTD_kokkos/FDTD_kokkos.h:31) that its driver fills on the
        device (perf-tests/kokkos_sample/kokkos_sample.cpp:82-108).
        ``torch.as_tensor(a, device="cuda")`` aliases the returned array.
        J arrays are FP64 on FP64 grids, FP32 otherwise, and stay at the same
        address (J becomes dense, which turns temporal blocking off); E and B
        hold the storage type (FP16: scale * value) and are valid until the
        next step.  The grid's work runs on ``torch_stream()``: do the
        writes there, or synchronize both ways.  From this call on, host
        mirrors are refreshed after every stepping or synchronizing call.
        torch brings its own HIP runtime: import torch and touch
        ``torch.cuda`` before the first grid is created (started after this
        library has launched work, it reports no device).  ``close()``
        frees the arrays whatever aliases them: tensors made from a
        DeviceArray must not be used after it."""
        try:
            comp = Component(component)
        except ValueError:
            raise LookupError("ERROR: Invalid field component") from None
        self._flush()
        ptr, eb, sc = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_double()
        _lib.check(self._L.fdtd_device_field(self._h, int(comp), ctypes.byref(ptr),
                                             ctypes.byref(eb), ctypes.byref(sc)))
        self._device = True
        self._invalidate()
        dtype = {8: np.float64, 4: np.float32, 2: np.float16}[eb.value]
        return DeviceArray(ptr.value, np.dtype(dtype), self.cells, sc.value, self)

    def stream_handle(self) -> int:
        """The grid's HIP stream (hipStream_t) as an integer."""
        s = ctypes.c_void_p()
        _lib.check(self._L.fdtd_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def torch_stream(self):
        """The grid's stream as a ``torch.cuda.ExternalStream``."""
        import torch

        return torch.cuda.ExternalStream(self.stream_handle())

    def _box(self, lo, hi):
        lo = self.begin_main if lo is None else lo
        hi = self.end_main if hi is None else hi
        return (ctypes.c_int * 3)(*[int(v) for v in lo]), (ctypes.c_int * 3)(*[int(v) for v in hi])

    def update_B(self, lo=None, hi=None) -> None:
        """The reference's protected ``update_B()`` (src/FDTD/FDTD.cpp:99-130):
        one B half step over the cells [lo, hi), by default the main box
        (``begin_main`` .. ``end_main``); neighbours wrap over the whole grid.
        Does not advance ``step``."""
        self._flush()
        a, b = self._box(lo, hi)
        _lib.check(self._L.fdtd_update_b_box(self._h, a, b))
        self._invalidate()

    def update_E(self, lo=None, hi=None) -> None:
        """The reference's protected ``update_E()`` (src/FDTD/FDTD.cpp:63-97,
        J paired as Jx->Ex, Jy->Ey, Jz->Ez like
        include/FDTD_kokkos/kokkos_functors.h:64-90): one E step with J and
        the point sources of the current step over [lo, hi), by default the
        main box.  Does not advance ``step``."""
        self._flush()
        a, b = self._box(lo, hi)
        _lib.check(self._L.fdtd_update_e_box(self._h, a, b))
        self._invalidate()

    def fields(self):
        """[Ex, Ey, Ez, Bx, By, Bz] as FP64 numpy arrays."""
        self._flush()
        return [self.download(c) for c in range(6)]

    def sum_squares(self):
        out = (ctypes.c_double * 6)()
        _lib.check(self._L.fdtd_sum_squares(self._h, out))
        return list(out)

    def max_abs(self):
        out = (ctypes.c_double * 6)()
        _lib.check(self._L.fdtd_max_abs(self._h, out))
        return list(out)

    def compare(self, ref: "FDTD"):
        """Per component (sum (t-r)^2, sum r^2, max|t-r|) against an FP64 grid."""
        out = (ctypes.c_double * 18)()
        _lib.check(self._L.fdtd_compare(self._h, ref._h, out))
        return [tuple(out[3 * c:3 * c + 3]) for c in range(6)]

    def add_plane_wave(self, m, amplitude, phase, polarization) -> None:
        """Add amplitude*e*cos(k.r+phase) to E and khat x E to B (device-side,
        Yee positions; k = 2 pi (m0/Ni, m1/Nj, m2/Nk) per cell)."""
        self._flush()
        mm = (ctypes.c_int * 3)(*[int(x) for x in m])
        pol = (ctypes.c_double * 3)(*[float(x) for x in polarization])
        _lib.check(self._L.fdtd_add_plane_wave(self._h, mm, float(amplitude), float(phase), pol))
        self._invalidate()

    def checksum(self):
        """Order-independent 64-bit checksums of the stored bits of Ex..Bz."""
        out = (ctypes.c_ulonglong * 6)()
        _lib.check(self._L.fdtd_checksum(self._h, out))
        return list(out)

    def set_scales(self, scales) -> None:
        arr = (ctypes.c_double * 6)(*[float(s) for s in scales])
        _lib.check(self._L.fdtd_set_scales(self._h, arr))

    def get_scales(self):
        arr = (ctypes.c_double * 6)()
        _lib.check(self._L.fdtd_get_scales(self._h, arr))
        return list(arr)

    def autoscale(self, target_log2: int = 12, magnitude_floor: float = 0.0) -> None:
        _lib.check(self._L.fdtd_autoscale(self._h, int(target_log2), float(magnitude_floor)))

    def set_autoscale_interval(self, interval: int, target_log2: int = 12) -> None:
        _lib.check(self._L.fdtd_set_autoscale_interval(self._h, int(interval), int(target_log2)))

    def zero_fields(self) -> None:
        """E, B, J := 0 and step := 0 on the device (sources are kept)."""
        for f in self._fields.values():
            f._host = None
            f._dirty = False
        _lib.check(self._L.fdtd_zero_fields(self._h))

    def set_time_blocking(self, L: int, V: int = 0, lanes: int = 0, TYg: int = 0,
                          kchunk: int = 0) -> None:
        """L steps per pass over HBM (L = 0: full-row sweep, 1: tiled single
        step); bit-identical results."""
        _lib.check(self._L.fdtd_set_time_blocking(self._h, int(L), int(V), int(lanes), int(TYg),
                                                  int(kchunk)))

    def set_time_blocking_strips(self, L: int, V: int = 0, R: int = 0, W: int = 0,
                                 kchunk: int = 0) -> None:
        """L steps per pass with the wavefront-strip kernel: tiles of 64*V
        columns by W*R rows, R rows per lane (0 = defaults); bit-identical
        results."""
        _lib.check(self._L.fdtd_set_time_blocking_strips(self._h, int(L), int(V), int(R), int(W),
                                                         int(kchunk)))

    def get_time_blocking_strips(self):
        """{L, V, R, W, kchunk}; R = W = 0 unless the strip kernel is set."""
        out = (ctypes.c_int * 5)()
        _lib.check(self._L.fdtd_get_time_blocking_strips(self._h, out))
        return dict(L=out[0], V=out[1], R=out[2], W=out[3], kchunk=out[4])

    def autotune_blocking(self, reps: int = 3):
        """Fastest of the full-row sweep and tiled/blocked passes (ms per step)."""
        out = (ctypes.c_int * 5)()
        ms = ctypes.c_float()
        _lib.check(self._L.fdtd_autotune_blocking(self._h, int(reps), out, ctypes.byref(ms)))
        cfg = dict(L=out[0], V=out[1], lanes=out[2], TYg=out[3], kchunk=out[4])
        strips = self.get_time_blocking_strips()
        cfg.update(R=strips["R"], W=strips["W"])
        return cfg, ms.value

    def get_time_blocking(self):
        out = (ctypes.c_int * 5)()
        _lib.check(self._L.fdtd_get_time_blocking(self._h, out))
        return dict(L=out[0], V=out[1], lanes=out[2], TYg=out[3], kchunk=out[4])

    def set_step_carry(self, enable: bool) -> None:
        """Carry the next call's first B half step across update_fields calls
        (on by default; bit-identical either way)."""
        _lib.check(self._L.fdtd_set_step_carry(self._h, int(bool(enable))))

    def set_tuning(self, V=0, TY=0, kchunk=0) -> None:
        _lib.check(self._L.fdtd_set_tuning(self._h, int(V), int(TY), int(kchunk)))

    def get_tuning(self):
        out = (ctypes.c_int * 3)()
        _lib.check(self._L.fdtd_get_tuning(self._h, out))
        return tuple(out)

    def autotune(self, reps: int = 5):
        out = (ctypes.c_int * 3)()
        ms = ctypes.c_float()
        _lib.check(self._L.fdtd_autotune(self._h, int(reps), out, ctypes.byref(ms)))
        return tuple(out), ms.value

    def time_update_fields_kernels(self, nsteps: int):
        """update_fields(nsteps) timed launch by launch: {kind: (ms, launches)} for
        'prologue' (B/2 without a carried step), 'steady' (passes or sweeps),
        'last' (the call's last pass or sweep) and 'remainder' (single sweeps
        after a blocked call's passes)."""
        self._flush()
        ms = (ctypes.c_double * 4)()
        cnt = (ctypes.c_int * 4)()
        _lib.check(self._L.fdtd_time_update_fields_kernels(self._h, int(nsteps), ms, cnt))
        self._invalidate()
        return {k: (ms[i], cnt[i]) for i, k in enumerate(("prologue", "steady", "last",
                                                          "remainder"))}

    def time_update_fields(self, nsteps: int, path: int = _lib.PATH_FUSED):
        """(ms_total, ms_main) of nsteps steps, HIP events on the grid's stream."""
        self._flush()
        a, b = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._L.fdtd_time_update_fields(self._h, int(nsteps), int(path),
                                                   ctypes.byref(a), ctypes.byref(b)))
        self._invalidate()
        return a.value, b.value


class FDTD_PML(FDTD):
    """The reference's PML subclass, ``FDTD_PML(Parameters, FP dt, FP pml_percent)``
    (include/FDTD/FDTD_PML.h:10-44, src/FDTD/FDTD_PML.cpp:205-365): the outer
    ``int(N*pml_percent)`` cells of every axis absorb outgoing waves; J acts
    on the main box only.  Same methods as FDTD, plus the layer's widths and
    split fields."""

    SPLITS = ("Exy", "Exz", "Eyx", "Eyz", "Ezx", "Ezy",
              "Bxy", "Bxz", "Byx", "Byz", "Bzx", "Bzy")  # FDTD_PML.h:12-13

    def __init__(self, parameters: Parameters, dt: float, pml_percent: float,
                 precision: Precision = Precision.FP64):
        self._open(parameters, dt, precision, pml_percent)
        # main box = interior, src/FDTD/FDTD_PML.cpp:259-268
        p = self.pml_widths
        self.begin_main = tuple(p)
        self.end_main = tuple(n - w for n, w in zip(self.shape, p))

    @property
    def pml_widths(self):
        w = (ctypes.c_int * 3)()
        pct = ctypes.c_double()
        _lib.check(self._L.fdtd_get_pml(self._h, w, ctypes.byref(pct)))
        return tuple(w)

    def set_pml_path(self, path: int, V: int = 0, TY: int = 0, kchunk: int = 0) -> None:
        """0: the fused PML sweep (default); 1: two in-place phase sweeps per step."""
        _lib.check(self._L.fdtd_set_pml_path(self._h, int(path), int(V), int(TY), int(kchunk)))

    def get_pml_path(self):
        out = (ctypes.c_int * 4)()
        _lib.check(self._L.fdtd_get_pml_path(self._h, out))
        return dict(path=out[0], V=out[1], TY=out[2], kchunk=out[3])

    def split(self, name) -> np.ndarray:
        """FP64 copy of one split field, by name ("Exy", ...) or index 0..11."""
        q = self.SPLITS.index(name) if isinstance(name, str) else int(name)
        self._flush()
        out = np.empty(self.cells, dtype=np.float64)
        _lib.check(self._L.fdtd_pml_download_split(self._h, q, _dptr(out), out.size))
        return out


def pml_profile(n: int, pml_percent: float, d: float, dt: float):
    """(sigma, decay, coef) of the layer along one axis (host-only, no GPU)."""
    L = _lib.load()
    out = [np.empty(n, dtype=np.float64) for _ in range(3)]
    _lib.check(L.fdtd_pml_profile(int(n), float(pml_percent), float(d), float(dt),
                                  *[_dptr(a) for a in out]))
    return tuple(out)


def set_tune_cache(path) -> None:
    """Keep fdtd_autotune_blocking's picks in the text file `path` (None: off)."""
    L = _lib.load()
    _lib.check(L.fdtd_set_tune_cache(None if path is None else str(path).encode()))


def calibration_copy(nbytes: int, iters: int = 10) -> float:
    """Average ms of one streaming device copy of nbytes."""
    ms = ctypes.c_float()
    _lib.check(_lib.load().fdtd_calibration_copy(int(nbytes), int(iters), ctypes.byref(ms)))
    return ms.value

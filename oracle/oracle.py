"""CPU oracle bindings -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  It wraps two shared libraries built by
oracle/Makefile:

* ``liboracle.so``: the from-scratch C restatement of the reference's Yee
  step (fdtd_oracle.c; E update per
  include/FDTD_kokkos/kokkos_functors.h:64-90, B update per :128-151, step
  order per src/FDTD_kokkos/FDTD_kokkos.cpp:91-104).
* ``_ref/libfdtd_ref.so`` (optional): the reference's own
  src/FDTD/FDTD.cpp compiled in place with its CMake flags, wrapped by
  ref_shim.cpp.  Present only where the reference checkout was available
  at build time.

Fields are numpy float64 arrays of N = Ni*Nj*Nk cells in the reference's
layout, index = i + j*Ni + k*Ni*Nj (src/FDTD/FDTD.cpp:66-78).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfdtd_ref.so")

C = 3e10              # include/Constants.h:9
PI = 3.14159265358    # include/Constants.h:14 (truncated)

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_coefficients.argtypes = [ctypes.c_double] * 4 + [_dp]
        L.oracle_update_B.argtypes = [ctypes.c_int] * 3 + [_dp] * 7
        L.oracle_update_E.argtypes = [ctypes.c_int] * 3 + [_dp] * 10
        L.oracle_update_fields.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 4
                                           + [ctypes.c_int] + [ctypes.POINTER(_dp)] * 3)
        L.oracle_waveform.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6
        L.oracle_waveform.restype = ctypes.c_double
        L.oracle_run_point_sources.argtypes = (
            [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_longlong, ctypes.c_int]
            + [ctypes.POINTER(_dp)] * 3 + [ctypes.c_int, _dp])
        L.oracle_sum_squares.argtypes = [_dp, ctypes.c_size_t]
        L.oracle_sum_squares.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    """Load the compiled reference class (oracle/_ref/libfdtd_ref.so)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference missing: {REF_SO}")
        L = ctypes.CDLL(REF_SO)
        L.ref_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 10
        L.ref_create.restype = ctypes.c_void_p
        L.ref_destroy.argtypes = [ctypes.c_void_p]
        L.ref_set.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong]
        L.ref_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong]
        L.ref_update_fields.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_zeroed_currents.argtypes = [ctypes.c_void_p]
        _ref = L
    return _ref


def coefficients(dx, dy, dz, dt):
    """{cEx, cEy, cEz, cBx, cBy, cBz, cur} in the order of src/FDTD/FDTD.cpp:43-53."""
    out = np.zeros(7)
    lib().oracle_coefficients(dx, dy, dz, dt, _ptr(out))
    return out


class OracleGrid:
    """A CPU grid advanced by the oracle: E = [Ex,Ey,Ez], B = [Bx,By,Bz], J = [Jx,Jy,Jz]."""

    def __init__(self, Ni, Nj, Nk, dx, dy, dz, dt):
        if Ni <= 0 or Nj <= 0 or Nk <= 0 or dt <= 0:
            raise ValueError("ERROR: invalid parameters")
        self.Ni, self.Nj, self.Nk = Ni, Nj, Nk
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        n = Ni * Nj * Nk
        self.E = [np.zeros(n) for _ in range(3)]
        self.B = [np.zeros(n) for _ in range(3)]
        self.J = [np.zeros(n) for _ in range(3)]
        self.step = 0

    def fields(self):
        """[Ex, Ey, Ez, Bx, By, Bz, Jx, Jy, Jz] (Component ordinals, include/Enums.h:5)."""
        return self.E + self.B + self.J

    def update_fields(self, nsteps=1):
        E = (_dp * 3)(*[_ptr(a) for a in self.E])
        B = (_dp * 3)(*[_ptr(a) for a in self.B])
        J = (_dp * 3)(*[_ptr(a) for a in self.J])
        lib().oracle_update_fields(self.Ni, self.Nj, self.Nk, self.dx, self.dy, self.dz,
                                   self.dt, nsteps, E, B, J)
        self.step += nsteps

    def update_B(self):
        c = coefficients(self.dx, self.dy, self.dz, self.dt)
        lib().oracle_update_B(self.Ni, self.Nj, self.Nk, _ptr(c), *[_ptr(a) for a in self.E],
                              *[_ptr(a) for a in self.B])

    def update_E(self):
        c = coefficients(self.dx, self.dy, self.dz, self.dt)
        lib().oracle_update_E(self.Ni, self.Nj, self.Nk, _ptr(c), *[_ptr(a) for a in self.E],
                              *[_ptr(a) for a in self.B], *[_ptr(a) for a in self.J])

    def run_point_sources(self, sources, nsteps):
        """Advance nsteps with point sources (list of dicts, see source_array)."""
        E = (_dp * 3)(*[_ptr(a) for a in self.E])
        B = (_dp * 3)(*[_ptr(a) for a in self.B])
        J = (_dp * 3)(*[_ptr(a) for a in self.J])
        src = source_array(sources)
        lib().oracle_run_point_sources(self.Ni, self.Nj, self.Nk, self.dx, self.dy, self.dz,
                                       self.dt, self.step, nsteps, E, B, J, len(sources),
                                       _ptr(src))
        self.step += nsteps

    def sum_squares(self):
        return [lib().oracle_sum_squares(_ptr(a), a.size) for a in self.E + self.B]


def source_array(sources):
    """Pack sources {i,j,k,comp,kind,amp,period,t0,sigma,phase} into the C layout."""
    rows = []
    for s in sources:
        rows.append([s["i"], s["j"], s["k"], s["comp"], s.get("kind", 1), s.get("amp", 1.0),
                     s.get("period", 1.0), s.get("t0", 0.0), s.get("sigma", 0.0),
                     s.get("phase", 0.0)])
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1, 10))


def waveform(kind, amp, period, t0, sigma, phase, t):
    return lib().oracle_waveform(kind, amp, period, t0, sigma, phase, t)


class RefGrid:
    """The reference's own FDTD_openmp::FDTD (compiled from its sources)."""

    def __init__(self, Ni, Nj, Nk, dx, dy, dz, dt, ax=0.0, bx=1.0, ay=0.0, by=1.0, az=0.0, bz=1.0):
        self.n = Ni * Nj * Nk
        self.h = ref_lib().ref_create(Ni, Nj, Nk, ax, bx, ay, by, az, bz, dx, dy, dz, dt)
        if not self.h:
            raise ValueError("ERROR: invalid parameters")

    def __del__(self):
        if getattr(self, "h", None):
            ref_lib().ref_destroy(self.h)
            self.h = None

    def set(self, comp, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        rc = ref_lib().ref_set(self.h, comp, _ptr(a), a.size)
        if rc != 0:
            raise ValueError(f"ref_set failed rc={rc}")

    def get(self, comp):
        out = np.empty(self.n)
        rc = ref_lib().ref_get(self.h, comp, _ptr(out), self.n)
        if rc != 0:
            raise ValueError(f"ref_get failed rc={rc}")
        return out

    def update_fields(self, nsteps=1):
        ref_lib().ref_update_fields(self.h, nsteps)

    def zeroed_currents(self):
        ref_lib().ref_zeroed_currents(self.h)

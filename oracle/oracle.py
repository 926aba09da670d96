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

C = 3e10              # include/Constants.h:6
PI = 3.14159265358    # include/Constants.h:11 (truncated)

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
        _ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_update_B_box.argtypes = [ctypes.c_int] * 3 + [_dp, _ip] + [_dp] * 6
        L.oracle_update_E_box.argtypes = [ctypes.c_int] * 3 + [_dp, _ip] + [_dp] * 9
        L.oracle_update_fields.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 4
                                           + [ctypes.c_int] + [ctypes.POINTER(_dp)] * 3)
        L.oracle_waveform.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6
        L.oracle_waveform.restype = ctypes.c_double
        L.oracle_run_point_sources.argtypes = (
            [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_longlong, ctypes.c_int]
            + [ctypes.POINTER(_dp)] * 3 + [ctypes.c_int, _dp])
        L.oracle_sum_squares.argtypes = [_dp, ctypes.c_size_t]
        L.oracle_sum_squares.restype = ctypes.c_double
        L.oracle_pml_sizes.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double,
                                                            ctypes.POINTER(ctypes.c_int)]
        L.oracle_pml_sigma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_pml_run.argtypes = (
            [ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [ctypes.c_longlong, ctypes.c_int]
            + [ctypes.POINTER(_dp)] * 5 + [ctypes.c_int, _dp, _dp])
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
        L.ref_update_B_box.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.ref_update_E_box.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.ref_set.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong]
        L.ref_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_longlong]
        L.ref_set_value.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                                    ctypes.c_double]
        L.ref_update_fields.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_zeroed_currents.argtypes = [ctypes.c_void_p]
        L.ref_pml_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 11
        L.ref_pml_create.restype = ctypes.c_void_p
        L.ref_pml_destroy.argtypes = [ctypes.c_void_p]
        L.ref_pml_update_fields.argtypes = [ctypes.c_void_p, ctypes.c_int]
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

    def update_B_box(self, lo, hi):
        """update_B over the cells [lo, hi) (lo, hi: (i, j, k))."""
        c = coefficients(self.dx, self.dy, self.dz, self.dt)
        box = (ctypes.c_int * 6)(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
        lib().oracle_update_B_box(self.Ni, self.Nj, self.Nk, _ptr(c), box,
                                  *[_ptr(a) for a in self.E], *[_ptr(a) for a in self.B])

    def update_E_box(self, lo, hi):
        """update_E over the cells [lo, hi)."""
        c = coefficients(self.dx, self.dy, self.dz, self.dt)
        box = (ctypes.c_int * 6)(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
        lib().oracle_update_E_box(self.Ni, self.Nj, self.Nk, _ptr(c), box,
                                  *[_ptr(a) for a in self.E], *[_ptr(a) for a in self.B],
                                  *[_ptr(a) for a in self.J])

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


def pml_sizes(Ni, Nj, Nk, pml_percent):
    """Layer widths (p_i, p_j, p_k), src/FDTD/FDTD_PML.cpp:254-256."""
    out = (ctypes.c_int * 3)()
    lib().oracle_pml_sizes(Ni, Nj, Nk, pml_percent, out)
    return tuple(out)


def pml_sigma(n, p, d):
    """Conductivity profile of one axis, src/FDTD/FDTD_PML.cpp:3-61,316-338."""
    out = np.zeros(n)
    lib().oracle_pml_sigma(n, p, d, _ptr(out))
    return out


class OraclePmlGrid(OracleGrid):
    """A CPU grid with the reference's PML layer (FDTD_PML, src/FDTD/FDTD_PML.cpp).

    ES = [Exy, Exz, Eyx, Eyz, Ezx, Ezy] and BS = [Bxy, Bxz, Byx, Byz, Bzx, Bzy]
    are the split fields (include/FDTD/FDTD_PML.h:12-13)."""

    def __init__(self, Ni, Nj, Nk, dx, dy, dz, dt, pml_percent):
        super().__init__(Ni, Nj, Nk, dx, dy, dz, dt)
        self.pml_percent = pml_percent
        n = Ni * Nj * Nk
        self.ES = [np.zeros(n) for _ in range(6)]
        self.BS = [np.zeros(n) for _ in range(6)]
        self._sigma = np.zeros(Ni + Nj + Nk)

    def _run(self, nsteps, sources):
        E = (_dp * 3)(*[_ptr(a) for a in self.E])
        B = (_dp * 3)(*[_ptr(a) for a in self.B])
        ES = (_dp * 6)(*[_ptr(a) for a in self.ES])
        BS = (_dp * 6)(*[_ptr(a) for a in self.BS])
        J = (_dp * 3)(*[_ptr(a) for a in self.J])
        src = source_array(sources) if sources else np.zeros(10)
        lib().oracle_pml_run(self.Ni, self.Nj, self.Nk, self.dx, self.dy, self.dz, self.dt,
                             self.pml_percent, self.step, nsteps, E, B, ES, BS, J,
                             len(sources) if sources else 0, _ptr(src), _ptr(self._sigma))
        self.step += nsteps

    def update_fields(self, nsteps=1):
        self._run(nsteps, None)

    def run_point_sources(self, sources, nsteps):
        self._run(nsteps, sources)

    def update_B(self):
        raise NotImplementedError("the PML grid steps only as a whole")

    update_E = update_B


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

    def set_value(self, comp, idx, value):
        rc = ref_lib().ref_set_value(self.h, comp, idx, value)
        if rc != 0:
            raise ValueError(f"ref_set_value failed rc={rc}")

    def update_fields(self, nsteps=1):
        ref_lib().ref_update_fields(self.h, nsteps)

    def zeroed_currents(self):
        ref_lib().ref_zeroed_currents(self.h)

    def update_B_box(self, lo, hi):
        """The class's protected update_B() over the cells [lo, hi)."""
        ref_lib().ref_update_B_box(self.h, (ctypes.c_int * 6)(lo[0], hi[0], lo[1], hi[1],
                                                               lo[2], hi[2]))

    def update_E_box(self, lo, hi):
        """The class's protected update_E() over [lo, hi) (Jx feeds all three
        E components there, src/FDTD/FDTD.cpp:85-93)."""
        ref_lib().ref_update_E_box(self.h, (ctypes.c_int * 6)(lo[0], hi[0], lo[1], hi[1],
                                                               lo[2], hi[2]))


class RefPmlGrid(RefGrid):
    """The reference's own FDTD_openmp::FDTD_PML (src/FDTD/FDTD_PML.cpp), run
    on one OpenMP thread (its layer updates share per-cell coefficients
    between threads)."""

    def __init__(self, Ni, Nj, Nk, dx, dy, dz, dt, pml_percent, ax=0.0, bx=1.0, ay=0.0,
                 by=1.0, az=0.0, bz=1.0):
        self.n = Ni * Nj * Nk
        self.h = ref_lib().ref_pml_create(Ni, Nj, Nk, ax, bx, ay, by, az, bz, dx, dy, dz, dt,
                                          pml_percent)
        if not self.h:
            raise ValueError("ERROR: invalid parameters")

    def __del__(self):
        if getattr(self, "h", None):
            ref_lib().ref_pml_destroy(self.h)
            self.h = None

    def update_fields(self, nsteps=1, one_thread=True):
        """one_thread=False runs the class on the whole OpenMP team, as the
        reference's drivers do (for timing only: the values may differ)."""
        if one_thread:
            ref_lib().ref_pml_update_fields(self.h, nsteps)
        else:
            ref_lib().ref_update_fields(self.h, nsteps)


# ---------------------------------------------------------------------------
# Independent numpy restatement (pins the C oracle's wrap and indexing).
#
# Arrays are viewed as [k, j, i] (index = i + j*Ni + k*Ni*Nj); the periodic
# neighbours of include/FDTD/FDTD.h:24-31 become np.roll along an axis.  The
# expression order is the reference's, so results are bit-identical to a
# correct C implementation (numpy does no FMA contraction).

def roll_update_fields(E, B, J, shape, coefs, nsteps=1):
    """nsteps of B/2, E, B/2 on float64 arrays (lists of 3 flat arrays, in place)."""
    Ni, Nj, Nk = shape
    cEx, cEy, cEz, cBx, cBy, cBz, cur = coefs
    v = lambda a: a.reshape(Nk, Nj, Ni)  # noqa: E731
    Ex, Ey, Ez = (v(a) for a in E)
    Bx, By, Bz = (v(a) for a in B)
    Jx, Jy, Jz = (v(a) for a in J)
    nxt = lambda a, ax: np.roll(a, -1, axis=ax)  # noqa: E731  a[.. +1 ..]
    prv = lambda a, ax: np.roll(a, 1, axis=ax)  # noqa: E731   a[.. -1 ..]
    I_, J_, K_ = 2, 1, 0

    def half_b():
        dbx = cBz * (nxt(Ey, K_) - Ey) - cBy * (nxt(Ez, J_) - Ez)
        dby = cBx * (nxt(Ez, I_) - Ez) - cBz * (nxt(Ex, K_) - Ex)
        dbz = cBy * (nxt(Ex, J_) - Ex) - cBx * (nxt(Ey, I_) - Ey)
        Bx[...] = Bx + dbx
        By[...] = By + dby
        Bz[...] = Bz + dbz

    for _ in range(nsteps):
        half_b()
        ex = Ex + ((cur * Jx + cEy * (Bz - prv(Bz, J_))) - cEz * (By - prv(By, K_)))
        ey = Ey + ((cur * Jy + cEz * (Bx - prv(Bx, K_))) - cEx * (Bz - prv(Bz, I_)))
        ez = Ez + ((cur * Jz + cEx * (By - prv(By, I_))) - cEy * (Bx - prv(Bx, J_)))
        Ex[...] = ex
        Ey[...] = ey
        Ez[...] = ez
        half_b()


# ---------------------------------------------------------------------------
# The reference drivers' spherical-wave harness (perf-tests/sample/sample.cpp:14-91)

_libm = None


def _m():
    global _libm
    if _libm is None:
        _libm = ctypes.CDLL("libm.so.6")
        for fn in ("sin", "cos"):
            getattr(_libm, fn).argtypes = [ctypes.c_double]
            getattr(_libm, fn).restype = ctypes.c_double
        _libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
        _libm.pow.restype = ctypes.c_double
    return _libm


def sample_current(x, y, z, t, T, Tx, Ty, Tz):
    """cur_func of sample.cpp:25-31, evaluated with the C library's libm."""
    m = _m()
    return (m.sin(2.0 * PI * t / T) * m.pow(m.cos(2.0 * PI * x / Tx), 2.0)
            * m.pow(m.cos(2.0 * PI * y / Ty), 2.0) * m.pow(m.cos(2.0 * PI * z / Tz), 2.0))


def sample_box(n):
    """Current box, parameters and dt of spherical_wave(n, it): sample.cpp:15-63."""
    period, mm, dt = 8, 4, 0.2
    T = float(period)
    Tx = Ty = Tz = float(mm) * C
    iterations = int(float(period) / dt)
    d = C
    boundary = float(n) / 2.0 * d
    ax = -boundary
    start = int(np.floor((-Tx / 4.0 - ax) / d))
    stop = int(np.floor((Tx / 4.0 - ax) / d))
    return dict(T=T, Tx=Tx, Ty=Ty, Tz=Tz, iterations=iterations, d=d, dt=dt, start=start,
                stop=stop, boundary=boundary)


def spherical_wave(n, it, pml_percent=None):
    """Run the harness on the oracle; returns [Ex..Bz] after `it` steps.

    J is written to Jx = Jy = Jz on the box before each of the first
    min(40, it) steps, then zeroed -- the pattern for which the reference's
    OpenMP and Kokkos E updates coincide.  With pml_percent the grid is the
    driver's FDTD_PML (sample.cpp:93-123)."""
    p = sample_box(n)
    if pml_percent is None:
        g = OracleGrid(n, n, n, p["d"], p["d"], p["d"], p["dt"])
    else:
        g = OraclePmlGrid(n, n, n, p["d"], p["d"], p["d"], p["dt"], pml_percent)
    cur_time = min(p["iterations"], it)
    for t in range(cur_time):
        for k in range(p["start"], p["stop"]):
            for j in range(p["start"], p["stop"]):
                for i in range(p["start"], p["stop"]):
                    idx = i + j * n + k * n * n
                    val = sample_current(float(i) * p["d"], float(j) * p["d"], float(k) * p["d"],
                                         float(t + 1) * p["dt"], p["T"], p["Tx"], p["Ty"], p["Tz"])
                    g.J[0][idx] = val
                    g.J[1][idx] = val
                    g.J[2][idx] = val
        g.update_fields(1)
    for a in g.J:
        a[:] = 0.0
    if it > cur_time:
        g.update_fields(it - cur_time)
    return g.E + g.B

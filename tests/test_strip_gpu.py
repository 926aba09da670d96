"""Temporal blocking with the wavefront-strip kernel (strip_sweep).

L fused steps per pass over HBM on tiles of 64*V columns by W*R rows must
give exactly the fields of L unblocked fused sweeps (bit-identical in every
precision, since the per-cell expressions are the same), with point
sources, remainders of n % L steps, several tiles and z chunks, and grids
smaller than a tile (ghost cells wrapping several times).  FP64 is also
held to the oracle.
"""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Parameters, Precision
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def grid(shape, dt, state, prec, sources=()):
    g = FDTD(Parameters(*shape), dt, prec)
    for c in range(6):
        g.upload(c, state[c])
    if sources:
        g.set_point_sources(list(sources))
    return g


def rel_l2(out, ref):
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(b ** 2)) for b in ref)
    return np.sqrt(num / den)


def sources_for(shape, dt):
    Ni, Nj, Nk = shape
    return [dict(i=3, j=Nj - 1, k=0, comp=2, kind=1, amp=1.0, period=16 * dt, t0=5 * dt,
                 sigma=6 * dt),
            dict(i=Ni - 1, j=0, k=Nk - 1, comp=0, kind=0, amp=0.3),
            dict(i=Ni // 2, j=Nj // 2, k=Nk // 2, comp=1, kind=1, amp=0.5, period=9 * dt,
                 t0=0.0, sigma=0.0)]


CASES = [
    # prec, L, V, R, W, kchunk
    (Precision.FP64, 1, 1, 2, 4, 5),
    (Precision.FP64, 2, 1, 2, 8, 0),
    (Precision.FP64, 3, 1, 2, 6, 7),
    (Precision.FP64, 4, 1, 2, 8, 4),
    (Precision.FP32, 4, 1, 2, 8, 0),
    (Precision.FP32, 1, 2, 2, 3, 0),
    (Precision.FP32, 2, 1, 4, 4, 6),
    (Precision.FP32, 2, 2, 2, 8, 0),
    (Precision.FP32, 3, 1, 4, 3, 0),
    (Precision.FP32, 3, 1, 2, 16, 9),
    (Precision.FP32, 4, 2, 2, 5, 3),
    (Precision.FP32, 4, 1, 4, 4, 0),
    (Precision.FP16, 2, 2, 2, 4, 0),
    (Precision.FP16, 3, 1, 4, 2, 5),
]


@pytest.mark.parametrize("prec,L,V,R,W,kchunk", CASES)
def test_strip_blocking_bit_exact(prec, L, V, R, W, kchunk):
    shape = (40, 36, 22)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1, factor=0.5)
    st = random_state(n, 40)
    src = sources_for(shape, dt)
    a = grid(shape, dt, st, prec, src)
    b = grid(shape, dt, st, prec, src)
    b.set_time_blocking_strips(L, V, R, W, kchunk)
    cfg = b.get_time_blocking_strips()
    print("strips", cfg, flush=True)
    assert (cfg["L"], cfg["V"], cfg["R"], cfg["W"]) == (L, V, R, W)
    for nsteps in (2 * L + 1, L, 3):
        a.update_fields(nsteps)
        b.update_fields(nsteps)
        assert a.checksum() == b.checksum(), nsteps
    if prec == Precision.FP64:
        r = O.OracleGrid(*shape, 1.0, 1.0, 1.0, dt)
        for c in range(6):
            r.fields()[c][:] = st[c]
        r.run_point_sources(src, 2 * L + 1 + L + 3)
        assert rel_l2(b.fields(), r.E + r.B) <= 1e-12


@pytest.mark.parametrize("shape", [(200, 130, 40), (130, 70, 65)], ids=lambda s: "x".join(map(str, s)))
def test_strip_blocking_many_tiles_and_chunks(shape):
    """Several tiles along x and y (partial last tiles) and z chunks."""
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 41)
    src = sources_for(shape, dt)
    a = grid(shape, dt, st, Precision.FP32, src)
    b = grid(shape, dt, st, Precision.FP32, src)
    b.set_time_blocking_strips(3, 1, 4, 4, 16)
    a.update_fields(7)
    b.update_fields(7)
    assert a.checksum() == b.checksum()
    b.set_time_blocking_strips(2, 2, 2, 8, 0)
    a.update_fields(5)
    b.update_fields(5)
    assert a.checksum() == b.checksum()


@pytest.mark.parametrize("storage", ["wide", "fp16"])
def test_strip_blocking_tiny_grids(storage):
    """Grids smaller than one tile: the x ghost columns and y ghost rows wrap
    several times around the grid (FP32 and FP64, or FP16 storage, whose
    strips read and write through plane buffer resources)."""
    for shape, V in [((4, 3, 5), 2), ((2, 2, 2), 2), ((6, 1, 3), 1), ((1, 5, 4), 1)]:
        n = int(np.prod(shape))
        dt = courant_dt(1, 1, 1)
        st = random_state(n, 42)
        if storage == "fp16":
            prec = Precision.FP16
            st = [x * 0.25 for x in st]
        else:
            prec = Precision.FP32 if V == 2 else Precision.FP64  # FP64 strips: V = 1
        a = grid(shape, dt, st, prec)
        b = grid(shape, dt, st, prec)
        b.set_time_blocking_strips(2, V, 2, 4, 1)
        a.update_fields(6)
        b.update_fields(6)
        assert a.checksum() == b.checksum(), shape


def test_strip_blocking_many_sources_on_one_plane():
    """Point sources sorted by plane: many on one plane and on the wrapped
    planes of a chunk's ghost range."""
    shape = (48, 40, 24)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 43)
    rng = np.random.default_rng(7)
    src = [dict(i=int(rng.integers(48)), j=int(rng.integers(40)), k=k, comp=int(rng.integers(3)),
                kind=1, amp=float(rng.uniform(0.1, 1)), period=float(rng.uniform(8, 30)) * dt,
                t0=0.0, sigma=0.0)
           for k in [0] * 20 + [23] * 5 + [7, 8, 9, 12]]
    a = grid(shape, dt, st, Precision.FP32, src)
    b = grid(shape, dt, st, Precision.FP32, src)
    b.set_time_blocking_strips(4, 1, 2, 8, 6)
    a.update_fields(9)
    b.update_fields(9)
    assert a.checksum() == b.checksum()


def test_strip_blocking_rejects_invalid_shapes():
    shape = (33, 20, 10)
    dt = courant_dt(1, 1, 1)
    g = grid(shape, dt, random_state(int(np.prod(shape)), 44), Precision.FP64)
    for args in [(2, 2, 2, 4, 0),    # V=2 with odd Ni
                 (2, 1, 4, 4, 0),    # FP64 builds V*R*8 <= 16 only
                 (2, 1, 3, 4, 0),    # R not built
                 (5, 1, 2, 4, 0),    # L > 4
                 (2, 1, 2, 1, 0),    # W*R - 2L < 1
                 (2, 1, 2, 17, 0)]:  # W > 16
        with pytest.raises(RuntimeError):
            g.set_time_blocking_strips(*args)
    g.set_time_blocking_strips(0)
    assert g.get_time_blocking_strips()["L"] == 0


@pytest.mark.parametrize("L,V,lanes,TYg", [(2, 2, 32, 12), (3, 1, 64, 8)])
def test_patch_blocking_fp16_bit_exact(L, V, lanes, TYg):
    """The ghost-zone patch kernel with FP16 storage: every level's fields
    are rounded to FP16 as between unblocked sweeps."""
    shape = (40, 36, 22)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1, factor=0.5)
    st = random_state(n, 45)
    src = sources_for(shape, dt)
    a = grid(shape, dt, st, Precision.FP16, src)
    b = grid(shape, dt, st, Precision.FP16, src)
    # patch passes start every call from the B/2 prologue (no carried half
    # step): compare with unblocked calls that do the same
    a.set_step_carry(False)
    b.set_time_blocking(L, V, lanes, TYg, 0)
    for nsteps in (2 * L + 1, L, 3):
        a.update_fields(nsteps)
        b.update_fields(nsteps)
        assert a.checksum() == b.checksum(), nsteps


@pytest.mark.parametrize("prec", [Precision.FP16, Precision.FP32])
def test_strip_blocking_carries_the_half_step(prec):
    """Strip runs carry the next half step across calls like the unblocked
    path: calls of L steps (and of L+1: a pass plus a fused remainder)
    equal one long unblocked call bitwise, FP16 included."""
    shape = (64, 40, 18)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = [x * 0.25 for x in random_state(n, 46)]
    src = sources_for(shape, dt)
    a = grid(shape, dt, st, prec, src)
    b = grid(shape, dt, st, prec, src)
    b.set_time_blocking_strips(3, 1, 2, 8, 0)
    for nsteps in (3, 3, 4, 3):
        b.update_fields(nsteps)
    a.update_fields(13)
    assert a.checksum() == b.checksum()


@pytest.mark.parametrize("kind", ["strips", "patch"])
def test_blocking_with_dense_currents_runs_single_sweeps(kind):
    """A dense J array (a whole J component uploaded) is read by single
    sweeps: with blocking set, such runs still equal unblocked ones, and
    blocking resumes once the dense J is dropped."""
    shape = (32, 20, 12)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 47)
    J = np.random.default_rng(5).standard_normal(n) * 1e-3
    a = grid(shape, dt, st, Precision.FP64)
    b = grid(shape, dt, st, Precision.FP64)
    if kind == "strips":
        b.set_time_blocking_strips(2, 1, 2, 8, 0)
    else:
        b.set_time_blocking(2, 1, 32, 12, 0)
    for g in (a, b):
        g.upload(7, J)  # dense JY
        g.update_fields(5)
    assert a.checksum() == b.checksum()
    for g in (a, b):
        g.zeroed_currents()
        g.update_fields(4)
    assert a.checksum() == b.checksum()


def test_strip_blocking_fp16_in_run_rescaling():
    """FP16 strip passes with in-run rescaling (a decision every 8 steps, two
    passes of L=4): a blocked run rescales the state in place between
    passes, where an unblocked run scales a sweep's output, so the two agree
    to FP16 rounding rather than bitwise.  The scales follow a growing
    point-source field, and the result stays within 5e-3 of FP64 and close to
    the unblocked rescaling run."""
    N = 40
    dt = courant_dt(1, 1, 1, factor=0.5)
    src = [dict(i=20, j=20, k=20, comp=2, kind=1, amp=1.0, period=20 * dt, t0=60 * dt,
                sigma=20 * dt)]

    def make(prec, blocking):
        g = FDTD(Parameters(N, N, N), dt, prec)
        g.set_point_sources(src)
        if prec == Precision.FP16:
            g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * dt * 1e-2)
            g.set_autoscale_interval(8, 10)
        if blocking:
            g.set_time_blocking_strips(*blocking)
        return g

    ref = make(Precision.FP64, None)
    ub = make(Precision.FP16, None)
    bl = make(Precision.FP16, (4, 1, 2, 8, 0))
    s0 = bl.get_scales()
    for g in (ref, ub, bl):
        g.update_fields(90)
    s1 = bl.get_scales()
    assert s1 != s0
    mx = bl.max_abs()
    for c in (0, 1, 2, 3, 4):  # Bz stays ~0 under a Jz source
        assert 2.0 ** 6 <= mx[c] * s1[c] <= 2.0 ** 14, (c, mx[c] * s1[c])

    err = bl.compare(ref)
    r_ref = np.sqrt(sum(e[0] for e in err) / sum(e[1] for e in err))
    r_ub = rel_l2(bl.fields(), ub.fields())
    print("fp16 blocked rescaling: relL2 vs FP64", r_ref, "vs unblocked FP16", r_ub, s0, s1)
    assert r_ref <= 5e-3
    assert r_ub <= 5e-3


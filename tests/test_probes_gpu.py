"""Probes and energy records, through the C-ABI.

Probes record FP64 values of single cells after every step on the device:
the fused sweeps and the strip passes write them from registers (the B of
a steady sweep is the first half step's sum, which the sweep does not
store), the three-sweep steps and PML grids gather them from memory.
FP64 probe rows are held bit-identical to the oracle's fields after each
step (the oracle stepped one step at a time), on every path, launch shape
and z chunking, across run boundaries and the carried half step.  FP32 and
FP16 rows equal the grid's own fields read back after each step (values as
stored, in physical units), FP16 with in-run rescaling included.

Energy records are the six sums of squares after every `every`-th step,
reduced on the device: bitwise fdtd_sum_squares of the same state, and
within 1e-12 (relative) of the oracle's sums.  Splitting a run at the
recorded steps leaves the fields bit-identical to an unsplit run.
"""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, FDTD_PML, Batch, Parameters, Precision
from fdtd_method_amd._lib import FDTDError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (40, 36, 22)


def sources_for(shape, dt):
    Ni, Nj, Nk = shape
    return [dict(i=3, j=Nj - 1, k=0, comp=2, kind=1, amp=1.0, period=16 * dt, t0=5 * dt,
                 sigma=6 * dt),
            dict(i=Ni - 1, j=0, k=Nk - 1, comp=0, kind=0, amp=0.3),
            dict(i=Ni // 2, j=Nj // 2, k=Nk // 2, comp=1, kind=1, amp=0.5, period=9 * dt,
                 t0=0.0, sigma=0.0)]


def constant_sources(shape):
    """Sources of constant value: the oracle's sources are then bitwise the
    device's (waveforms use the device's sin/exp, which may differ from the
    host's libm in the last bit)."""
    Ni, Nj, Nk = shape
    return [dict(i=3, j=Nj - 1, k=0, comp=2, kind=0, amp=1.0),
            dict(i=Ni - 1, j=0, k=Nk - 1, comp=0, kind=0, amp=0.3),
            dict(i=Ni // 2, j=Nj // 2, k=Nk // 2, comp=1, kind=0, amp=-0.5)]


def probe_cells(shape, seed, n=64):
    """Random cells of random components plus the corners and the planes
    whose plane keys wrap (E at k = 0 and B at k = Nk-1 share key 0), and
    one cell listed twice."""
    Ni, Nj, Nk = shape
    rng = np.random.default_rng(seed)
    cells = [(int(rng.integers(Ni)), int(rng.integers(Nj)), int(rng.integers(Nk)),
              int(rng.integers(6))) for _ in range(n)]
    for c in range(6):
        cells += [(0, 0, 0, c), (Ni - 1, Nj - 1, Nk - 1, c), (Ni // 2, 0, Nk - 1, c),
                  (1, Nj - 1, 0, c)]
    cells.append(cells[0])
    return cells


def oracle_rows(shape, dt, state, src, cells, nsteps, pml=None):
    r = (O.OracleGrid(*shape, 1.0, 1.0, 1.0, dt) if pml is None
         else O.OraclePmlGrid(*shape, 1.0, 1.0, 1.0, dt, pml))
    for c in range(6):
        r.fields()[c][:] = state[c]
    Ni, Nj, _ = shape
    idx = [i + j * Ni + k * Ni * Nj for i, j, k, _ in cells]
    rows = []
    for _ in range(nsteps):
        r.run_point_sources(src, 1)
        f = r.fields()
        rows.append([f[c][q] for q, (_, _, _, c) in zip(idx, cells)])
    return np.array(rows), r


def make(shape, dt, state, prec, src, cls=FDTD, pml=None):
    p = Parameters(*shape)
    g = cls(p, dt, prec) if pml is None else cls(p, dt, pml, prec)
    for c in range(6):
        g.upload(c, state[c])
    g.set_point_sources(src)
    return g


# blocking: None = single fused sweeps; ("tune", V, TY, kchunk) = single
# sweeps of that launch shape; (L, V, R, W, kchunk) = strip passes
PATHS = [None, ("tune", 1, 4, 4), ("tune", 2, 3, 5), ("tune", 4, 2, 0),
         (4, 1, 2, 8, 0), (3, 1, 2, 6, 7), (2, 1, 2, 4, 5), (4, 1, 2, 8, 4), "3sweep"]


def set_path(g, path):
    g.set_time_blocking_strips(0)
    if path is None or path == "3sweep":
        return
    if path[0] == "tune":
        g.set_tuning(*path[1:])
    else:
        g.set_time_blocking_strips(*path)


def step(g, path, n):
    if path == "3sweep":
        g.update_fields_3sweep(n)
    else:
        g.update_fields(n)


@pytest.mark.parametrize("path", PATHS, ids=str)
def test_fp64_probes_equal_oracle_every_step(path):
    dt = courant_dt(1, 1, 1, factor=0.5)
    n = int(np.prod(SHAPE))
    st = random_state(n, 60)
    src = constant_sources(SHAPE)
    cells = probe_cells(SHAPE, 61)
    calls = (9, 4, 1, 7)  # passes with remainders, whole passes, single steps
    expect, _ = oracle_rows(SHAPE, dt, st, src, cells, sum(calls))
    g = make(SHAPE, dt, st, Precision.FP64, src)
    set_path(g, path)
    g.set_probes(cells, capacity=sum(calls))
    for k in calls:
        step(g, path, k)
    assert g.probe_rows() == sum(calls)
    rows = g.read_probes()
    assert not np.isnan(rows).any()
    bad = np.argwhere(rows != expect)
    assert bad.size == 0, f"{len(bad)} values differ, first (step, probe) {bad[:5].tolist()}"
    # the read-back fields are the oracle's too (the probes did not change the run)
    g.close()


def test_fp64_probes_pml_gathered_every_step():
    shape, pml = (24, 20, 16), 0.25
    dt = courant_dt(1, 1, 1, factor=0.5)
    n = int(np.prod(shape))
    st = random_state(n, 67)
    src = constant_sources(shape)
    cells = probe_cells(shape, 62, n=32)
    expect, r = oracle_rows(shape, dt, st, src, cells, 12, pml=pml)
    g = make(shape, dt, st, Precision.FP64, src, cls=FDTD_PML, pml=pml)
    g.set_probes(cells, capacity=12)
    g.update_fields(5)
    g.update_fields(7)
    rows = g.read_probes()
    assert np.array_equal(rows, expect)
    for c in range(6):
        assert np.array_equal(g.download(c), r.fields()[c]), c
    g.close()


def twin_grids(prec, path, rescale=False):
    shape = (64, 40, 18)
    dt = courant_dt(1, 1, 1)
    n = int(np.prod(shape))
    st = [x * 0.25 for x in random_state(n, 63)]
    src = sources_for(shape, dt)
    grids = []
    for _ in range(2):
        g = make(shape, dt, st, prec, src)
        if prec == Precision.FP16:
            # power-of-two scales other than 1; with rescaling, the first
            # decision moves every scale down by about 4
            g.autoscale(target_log2=12 if rescale else 10)
            if rescale:
                g.set_autoscale_interval(4, 10)
        set_path(g, path)
        grids.append(g)
    cells = probe_cells(shape, 64)
    Ni, Nj, _ = shape
    idx = np.array([i + j * Ni + k * Ni * Nj for i, j, k, _ in cells])
    comp = np.array([c for *_, c in cells])
    return grids, cells, idx, comp


def at_cells(g, idx, comp):
    f = [g.download(c) for c in range(6)]
    return np.array([f[c][q] for q, c in zip(idx, comp)])


@pytest.mark.parametrize("prec,path", [(Precision.FP32, None), (Precision.FP32, (4, 1, 2, 8, 0)),
                                       (Precision.FP32, ("tune", 8, 2, 0)),
                                       (Precision.FP32, "3sweep"),
                                       (Precision.FP16, None), (Precision.FP16, (3, 1, 4, 4, 0))],
                         ids=str)
def test_probes_are_the_stored_values(prec, path):
    """Rows of one long run equal the fields read back after stepping a twin
    grid one call at a time (bit-identical runs: the carried half step)."""
    (a, b), cells, idx, comp = twin_grids(prec, path)
    nsteps = 14
    a.set_probes(cells, capacity=nsteps)
    step(a, path, nsteps)
    expect = []
    for _ in range(nsteps):
        step(b, path, 1)
        expect.append(at_cells(b, idx, comp))
    assert a.checksum() == b.checksum()
    assert np.array_equal(a.read_probes(), np.array(expect))
    a.close()
    b.close()


@pytest.mark.parametrize("path", [None, (3, 1, 4, 4, 0)], ids=str)
def test_fp16_probes_through_in_run_rescaling(path):
    """FP16 rescaled every 4 steps inside the run: the probes do not change
    the run, the last row is the final state, and every row (rescaling
    sweeps and passes included) agrees with an FP32 run to FP16 resolution,
    so no row was scaled by a stale scale (a power of two off)."""
    (a, b), cells, idx, comp = twin_grids(Precision.FP16, path, rescale=True)
    (f32, _), *_ = twin_grids(Precision.FP32, path)
    nsteps = 14
    s0 = a.get_scales()
    for g in (a, f32):
        g.set_probes(cells, capacity=nsteps)
        step(g, path, nsteps)
    step(b, path, nsteps)
    assert a.get_scales() != s0  # the rescaling ran
    assert a.checksum() == b.checksum()
    rows, ref = a.read_probes(), f32.read_probes()
    assert np.array_equal(rows[-1], at_cells(b, idx, comp))
    # FP16 keeps about 11 bits of each component's largest value
    tol = 2e-2 * np.array(f32.max_abs())[comp]
    assert np.all(np.abs(rows - ref) <= tol), (np.abs(rows - ref) / tol).max()
    big = np.abs(ref) > 0.2 * tol / 2e-2  # rows a stale power-of-two scale would move
    assert big.sum() > rows.size // 8
    for g in (a, b, f32):
        g.close()


@pytest.mark.parametrize("path", [None, (4, 1, 2, 8, 0), (3, 1, 2, 6, 7), "3sweep"], ids=str)
def test_energy_records(path):
    dt = courant_dt(1, 1, 1, factor=0.5)
    n = int(np.prod(SHAPE))
    st = random_state(n, 65)
    src = constant_sources(SHAPE)
    r = O.OracleGrid(*SHAPE, 1.0, 1.0, 1.0, dt)
    for c in range(6):
        r.fields()[c][:] = st[c]
    every, total = 5, 23
    oracle_sums = []
    for s in range(1, total + 1):
        r.run_point_sources(src, 1)
        if s % every == 0:
            oracle_sums.append(r.sum_squares())
    a = make(SHAPE, dt, st, Precision.FP64, src)
    b = make(SHAPE, dt, st, Precision.FP64, src)
    for g in (a, b):
        set_path(g, path)
    a.set_energy_interval(every, capacity=8)
    step(a, path, 13)
    step(a, path, total - 13)
    step(b, path, total)
    assert a.checksum() == b.checksum()  # the split run equals the unsplit one
    steps, sums = a.energy_records()
    assert steps == [5, 10, 15, 20]
    ref = np.array(oracle_sums)
    assert np.all(np.abs(sums - ref) <= 1e-12 * np.abs(ref)), np.abs(sums / ref - 1).max()
    # bitwise the on-demand reduction of the same state
    c = make(SHAPE, dt, st, Precision.FP64, src)
    step(c, path, 20)
    assert np.array_equal(sums[-1], np.array(c.sum_squares()))
    for g in (a, b, c):
        g.close()


def test_record_capacity_and_rewind():
    dt = courant_dt(1, 1, 1)
    n = int(np.prod(SHAPE))
    g = make(SHAPE, dt, random_state(n, 66), Precision.FP32, sources_for(SHAPE, dt))
    g.set_probes([(1, 2, 3, 0), (4, 5, 6, 5)], capacity=6)
    g.set_energy_interval(2, capacity=2)
    g.update_fields(4)
    with pytest.raises(FDTDError):
        g.update_fields(3)  # 7 rows > 6: nothing runs
    assert g.step == 4 and g.probe_rows() == 4
    with pytest.raises(FDTDError):
        g.update_fields(2)  # a third energy record
    g.rewind_records()
    g.update_fields(4)
    assert g.probe_rows() == 4
    steps, _ = g.energy_records()
    assert steps == [6, 8]
    g.set_probes([], capacity=0)
    g.set_energy_interval(0)
    g.update_fields(10)
    assert g.probe_rows() == 0
    g.close()


def test_batches_refuse_recording_grids():
    dt = courant_dt(1, 1, 1)
    grids = [FDTD(Parameters(32, 16, 8), dt, Precision.FP32) for _ in range(2)]
    grids[1].set_probes([(0, 0, 0, 2)], capacity=4)
    with pytest.raises(FDTDError):
        batch = Batch(grids)
        batch.update_fields(2)
    for g in grids:
        g.close()


def test_probes_and_energy_together_on_pml_and_fp16_strips():
    """Both records at once: on a PML grid (probes gathered after single-step
    runs, the run split at the energy records) FP64 rows equal the oracle's
    fields bitwise and the energies its sums to 1e-12; on FP16 strip passes
    with in-run rescaling, split at the energy records, the last row is the
    final state and the records fall on the requested steps."""
    shape, pml = (24, 20, 16), 0.25
    dt = courant_dt(1, 1, 1, factor=0.5)
    n = int(np.prod(shape))
    st = random_state(n, 68)
    src = constant_sources(shape)
    cells = probe_cells(shape, 69, n=24)
    expect, _ = oracle_rows(shape, dt, st, src, cells, 11, pml=pml)
    r = O.OraclePmlGrid(*shape, 1.0, 1.0, 1.0, dt, pml)
    for c in range(6):
        r.fields()[c][:] = st[c]
    sums = []
    for s in range(1, 12):
        r.run_point_sources(src, 1)
        if s % 3 == 0:
            sums.append(r.sum_squares())
    g = make(shape, dt, st, Precision.FP64, src, cls=FDTD_PML, pml=pml)
    g.set_probes(cells, capacity=11)
    g.set_energy_interval(3, capacity=3)
    g.update_fields(4)
    g.update_fields(7)
    assert np.array_equal(g.read_probes(), expect)
    steps, got = g.energy_records()
    assert steps == [3, 6, 9]
    ref = np.array(sums)
    assert np.all(np.abs(got - ref) <= 1e-12 * np.abs(ref))
    g.close()

    (a, b), cells, idx, comp = twin_grids(Precision.FP16, (4, 1, 2, 8, 0), rescale=True)
    b.close()
    a.set_probes(cells, capacity=18)
    a.set_energy_interval(6, capacity=3)
    a.update_fields(18)
    rows = a.read_probes()
    assert not np.isnan(rows).any()
    assert np.array_equal(rows[-1], at_cells(a, idx, comp))
    steps, _ = a.energy_records()
    assert steps == [6, 12, 18]
    a.close()

"""Property-based parity (hypothesis, derandomized): random grid shapes,
launch shapes, step counts, spacings, layer widths and current patterns, the
FP64 fused path bit-identical to the oracle each time; and random temporal
blockings (strip passes) bit-identical to unblocked steps in every
precision.  Complements the hand-picked cases of test_parity_gpu.py /
test_pml_gpu.py / test_strip_gpu.py with shapes nobody chose (ragged tiles,
chunk boundaries on the last plane, one-cell axes)."""
import os

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from conftest import courant_dt
from fdtd_method_amd import FDTD, FDTD_PML, Parameters, Precision, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# FDTD_PROPERTY_EXAMPLES=N runs N cases per property; FDTD_PROPERTY_RANDOM=1
# draws new random cases on every run instead of the fixed derandomized set.
SETTINGS = settings(max_examples=int(os.environ.get("FDTD_PROPERTY_EXAMPLES", "40")),
                    deadline=None,
                    derandomize=os.environ.get("FDTD_PROPERTY_RANDOM", "0") != "1",
                    database=None,
                    suppress_health_check=[HealthCheck.too_slow, HealthCheck.filter_too_much])


@st.composite
def cases(draw, pml=False):
    Ni = draw(st.integers(1, 40))
    Nj = draw(st.integers(1, 24))
    Nk = draw(st.integers(1, 20))
    spacing = tuple(draw(st.sampled_from([0.7, 1.0, 1.3])) for _ in range(3))
    steps = draw(st.integers(1, 6))
    pct = draw(st.sampled_from([0.0, 0.1, 0.25, 0.5])) if pml else None
    V = draw(st.sampled_from([1, 2, 4]))
    TY = draw(st.integers(1, 8))
    kchunk = draw(st.integers(1, 20))
    seed = draw(st.integers(0, 2 ** 31 - 1))
    return (Ni, Nj, Nk), spacing, steps, pct, (V, TY, kchunk), seed


def run_case(shape, spacing, steps, pct, tune, seed):
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    rng = np.random.default_rng(seed)
    state = [rng.standard_normal(n) for _ in range(6)]
    J = np.zeros(n)
    J[rng.integers(0, n, size=min(n, 5))] = rng.standard_normal(min(n, 5))
    params = Parameters(*shape, dx=spacing[0], dy=spacing[1], dz=spacing[2])
    g = FDTD(params, dt) if pct is None else FDTD_PML(params, dt, pct)
    try:
        g.set_tuning(*tune)
    except _lib.FDTDError:
        pass  # shape not valid for this grid: keep the default
    for c in range(6):
        g.upload(c, state[c])
    for c in range(3):
        g.upload(6 + c, J)
    r = (O.OracleGrid(*shape, *spacing, dt) if pct is None
         else O.OraclePmlGrid(*shape, *spacing, dt, pct))
    for c in range(6):
        r.fields()[c][:] = state[c]
    for c in range(3):
        r.J[c][:] = J
    try:
        used = g.get_tuning()
    except _lib.FDTDError:
        used = "three-sweep"  # no fused launch shape fits this grid
    print(f"case shape={shape} spacing={spacing} steps={steps} pml={pct} tune={tune} "
          f"used={used} seed={seed}", flush=True)
    g.update_fields(steps)
    r.update_fields(steps)
    for c in range(6):
        assert np.array_equal(g.download(c), r.fields()[c]), f"component {c}"
    if pct is not None:
        for q in range(12):
            assert np.array_equal(g.split(q), (r.ES + r.BS)[q]), f"split {q}"


@SETTINGS
@given(cases())
def test_random_periodic_grids_bit_exact(case):
    run_case(*case)


@SETTINGS
@given(cases(pml=True))
def test_random_pml_grids_bit_exact(case):
    run_case(*case)


@st.composite
def strip_cases(draw):
    Ni = draw(st.integers(1, 80))
    Nj = draw(st.integers(1, 40))
    Nk = draw(st.integers(1, 20))
    prec = draw(st.sampled_from([Precision.FP64, Precision.FP32, Precision.FP16]))
    L = draw(st.integers(1, 4))
    V, R = draw(st.sampled_from([(1, 2)] if prec == Precision.FP64 else [(1, 2), (1, 4), (2, 2)]))
    W = draw(st.integers(1, 16))
    kchunk = draw(st.sampled_from([0, 1, 3, 7, 20]))
    calls = draw(st.lists(st.integers(1, 9), min_size=1, max_size=3))
    nsrc = draw(st.integers(0, 4))
    seed = draw(st.integers(0, 2 ** 31 - 1))
    return (Ni, Nj, Nk), prec, (L, V, R, W, kchunk), calls, nsrc, seed


@SETTINGS
@given(strip_cases())
def test_random_strip_blocking_bit_exact(case):
    """Temporally blocked strip passes on random shapes, precisions,
    blockings, call sequences and point sources: bitwise the unblocked
    fused steps (shapes the kernel rejects are skipped)."""
    shape, prec, blocking, calls, nsrc, seed = case
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    rng = np.random.default_rng(seed)
    state = [rng.standard_normal(n) * 0.25 for _ in range(6)]
    src = [dict(i=int(rng.integers(shape[0])), j=int(rng.integers(shape[1])),
                k=int(rng.integers(shape[2])), comp=int(rng.integers(3)), kind=int(rng.integers(2)),
                amp=float(rng.uniform(0.1, 1.0)), period=float(rng.uniform(8, 30)) * dt,
                t0=0.0, sigma=0.0) for _ in range(nsrc)]
    grids = []
    for blocked in (False, True):
        g = FDTD(Parameters(*shape), dt, prec)
        for c in range(6):
            g.upload(c, state[c])
        if src:
            g.set_point_sources(src)
        grids.append(g)
    try:
        grids[1].set_time_blocking_strips(*blocking)
    except _lib.FDTDError:
        return  # not a valid strip shape for this grid
    print(f"strip case shape={shape} prec={prec.name} blocking={grids[1].get_time_blocking_strips()} "
          f"calls={calls} nsrc={nsrc} seed={seed}", flush=True)
    for nsteps in calls:
        for g in grids:
            g.update_fields(nsteps)
        assert grids[0].checksum() == grids[1].checksum(), nsteps
    for g in grids:
        g.close()

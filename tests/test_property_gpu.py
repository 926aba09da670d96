"""Property-based parity (hypothesis, derandomized): random grid shapes,
launch shapes, step counts, spacings, layer widths and current patterns, the
FP64 fused path bit-identical to the oracle each time.  Complements the
hand-picked cases of test_parity_gpu.py / test_pml_gpu.py with shapes nobody
chose (ragged tiles, chunk boundaries on the last plane, one-cell axes)."""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

from conftest import courant_dt
from fdtd_method_amd import FDTD, FDTD_PML, Parameters, Precision, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=40, deadline=None, derandomize=True,
                    suppress_health_check=[HealthCheck.too_slow])


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

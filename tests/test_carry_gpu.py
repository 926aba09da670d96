"""The carried half step (fdtd_set_step_carry): consecutive update_fields
calls skip the B/2 prologue because the previous call's final sweep stored
(B_n + d) + d.  Checked: one-step loops with per-step currents (the
reference drivers' pattern, perf-tests/sample/sample.cpp:66-87) are
bit-identical to the oracle and to runs without carrying, and to one long
call; host writes to E/B, scale changes and batch runs drop the carried
step correctly."""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Batch, Parameters, Precision
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def grid(shape, dt, precision=Precision.FP64, state=None, carry=True):
    g = FDTD(Parameters(*shape), dt, precision)
    g.set_step_carry(carry)
    if state is not None:
        for c in range(6):
            g.upload(c, state[c])
    return g


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_one_step_loop_with_changing_currents(precision):
    shape = (24, 20, 16)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 31)
    a = grid(shape, dt, precision, st, carry=True)
    b = grid(shape, dt, precision, st, carry=False)
    r = O.OracleGrid(*shape, 1, 1, 1, dt)
    for c in range(6):
        r.fields()[c][:] = st[c]
    cells = [5 + 6 * 24 + 7 * 480, 100, n - 3]
    rng = np.random.default_rng(5)
    for t in range(7):
        vals = rng.standard_normal(len(cells))
        for g in (a, b):
            for q, idx in enumerate(cells):
                for comp in (6, 7, 8):
                    g.get_field(comp)[idx] = vals[q]
            g.update_fields()
        for q, idx in enumerate(cells):
            for comp in range(3):
                r.J[comp][idx] = vals[q]
        r.update_fields(1)
    for c in range(6):
        assert np.array_equal(a.download(c), b.download(c)), f"carry vs none, component {c}"
        if precision == Precision.FP64:
            assert np.array_equal(a.download(c), r.fields()[c]), f"oracle, component {c}"


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32, Precision.FP16])
def test_single_calls_equal_one_long_call(precision):
    """With the carried step, 6 calls of one step == one call of 6 steps
    bitwise in every precision (the carried value is the steady sweep's)."""
    shape = (32, 12, 10)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = [a * 0.25 for a in random_state(n, 8)]
    a = grid(shape, dt, precision, st)
    b = grid(shape, dt, precision, st)
    for _ in range(6):
        a.update_fields(1)
    b.update_fields(6)
    assert a.checksum() == b.checksum()


def test_host_writes_scales_and_batches_drop_the_carried_step():
    shape = (20, 16, 12)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 12)
    g = grid(shape, dt, Precision.FP64, st)
    r = O.OracleGrid(*shape, 1, 1, 1, dt)
    for c in range(6):
        r.fields()[c][:] = st[c]
    g.update_fields(2)
    r.update_fields(2)
    # host write to E and B between calls
    ex = g.download(0)
    ex[::7] = 0.25
    g.upload(0, ex)
    r.E[0][::7] = 0.25
    bz = g.download(5)
    bz[3::11] = -1.5
    g.upload(5, bz)
    r.B[2][3::11] = -1.5
    g.update_fields(1)
    r.update_fields(1)
    g.update_fields(3)
    r.update_fields(3)
    for c in range(6):
        assert np.array_equal(g.download(c), r.fields()[c]), f"component {c}"
    # a batch run after carrying calls starts from the grids' own state
    h = grid(shape, dt, Precision.FP64, st)
    h.update_fields(6)   # same step count as g (a batch requirement)
    Batch([h, g]).update_fields(2)
    r2 = O.OracleGrid(*shape, 1, 1, 1, dt)
    for c in range(6):
        r2.fields()[c][:] = st[c]
    r2.update_fields(8)
    r.update_fields(2)
    for c in range(6):
        assert np.array_equal(h.download(c), r2.fields()[c])
        assert np.array_equal(g.download(c), r.fields()[c])
    # FP16 scale change between calls
    f = grid(shape, dt, Precision.FP16, [a * 1e-3 for a in st])
    f.update_fields(1)
    f.set_scales([2.0 ** 12] * 6)
    f.update_fields(2)
    ref = grid(shape, dt, Precision.FP64, [a * 1e-3 for a in st])
    ref.update_fields(3)
    err = f.compare(ref)
    rel = np.sqrt(sum(e[0] for e in err) / sum(e[1] for e in err))
    print("fp16 after a scale change: relL2 vs FP64", rel)
    assert rel < 5e-3

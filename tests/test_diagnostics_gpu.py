"""Non-finite check and the autotuner cache, through the C-ABI.

* fdtd_count_nonfinite counts NaN/Inf values per component as stored: a
  NaN and an Inf forced into the fields from the host are counted exactly,
  spread to the neighbours the curl couples them to after a step, and an
  FP16 overflow (a storage scale too large for the values) counts every
  cell.
* fdtd_set_tune_cache: the autotuner's pick is written once per shape and
  precision; a second grid of that shape takes it without timing (no new
  line, the recorded ms), and steps bit-identically to the first.  A
  garbled line and a pick that does not fit the grid are ignored.
"""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Parameters, Precision, set_tune_cache

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32, Precision.FP16])
def test_nonfinite_count_host_forced(prec):
    shape = (32, 24, 16)
    Ni, Nj, Nk = shape
    n = Ni * Nj * Nk
    g = FDTD(Parameters(*shape), courant_dt(1, 1, 1), prec)
    g.set_time_blocking_strips(0)
    st = [0.01 * x for x in random_state(n, 70)]
    c0 = 5 + 7 * Ni + 9 * Ni * Nj
    st[0][c0] = np.nan        # Ex
    st[5][c0 + 3] = np.inf    # Bz
    st[5][c0 + 4] = -np.inf
    for c in range(6):
        g.upload(c, st[c])
    assert g.count_nonfinite() == [1, 0, 0, 0, 0, 2]
    g.update_fields(1)
    counts = g.count_nonfinite()
    print("after one step", counts, flush=True)
    # the NaN reaches Ex's own cell and the B components its curl feeds;
    # the infinities reach E
    assert counts[0] >= 1 and counts[4] >= 1 and counts[5] >= 2
    assert counts[1] >= 1
    assert np.isnan(g.sum_squares()).any() or np.isinf(g.sum_squares()).any()
    g.zero_fields()
    assert g.count_nonfinite() == [0] * 6
    g.close()


def test_nonfinite_count_fp16_overflow():
    shape = (32, 24, 16)
    n = int(np.prod(shape))
    g = FDTD(Parameters(*shape), courant_dt(1, 1, 1), Precision.FP16)
    g.set_scales([2.0 ** 20] * 6)   # values of 1 stored as 2^20 > 65504
    g.upload(2, np.ones(n))
    assert g.count_nonfinite() == [0, 0, n, 0, 0, 0]
    g.close()


def test_autotune_cache(tmp_path):
    path = tmp_path / "tune.txt"
    shape = (128, 64, 32)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 71)
    set_tune_cache(path)
    try:
        grids = []
        for _ in range(2):
            g = FDTD(Parameters(*shape), dt, Precision.FP32)
            for c in range(6):
                g.upload(c, st[c])
            grids.append(g)
        a, b = grids
        cfg, ms = a.autotune_blocking(reps=2)
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and "|" in lines[0], lines
        cfg2, ms2 = b.autotune_blocking(reps=2)
        assert cfg2 == cfg
        assert ms2 == pytest.approx(ms, rel=1e-5)
        assert len(path.read_text().splitlines()) == 1  # taken from the file, not timed
        a.update_fields(9)
        b.update_fields(9)
        assert a.checksum() == b.checksum()
        # a garbled line and a pick that does not fit the grid are ignored
        key = lines[0].split("|")[0].strip()
        with open(path, "a") as f:
            f.write("garbage without a bar\n")
            f.write(f"{key} | 4 1 64 999 0 2 0.001\n")  # W = 499.5: invalid
        c = FDTD(Parameters(*shape), dt, Precision.FP32)
        cfg3, ms3 = c.autotune_blocking(reps=2)
        assert ms3 != pytest.approx(0.001)
        assert len(path.read_text().splitlines()) == 4  # timed again and recorded
        for g in (a, b, c):
            g.close()
    finally:
        set_tune_cache(None)

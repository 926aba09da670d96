"""The allocation-free diagnostic entry point (fdtd_sum_squares_noalloc),
called through the C-ABI: its sums come from a fixed device array instead
of the grid's own partials buffer, and must equal fdtd_sum_squares bitwise,
on consecutive calls whose second grid is smaller (the shared array keeps
the first call's contents, which the second must not read)."""
import ctypes

import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Parameters, Precision, _lib

pytestmark = pytest.mark.gpu


def grid(shape, prec, seed):
    g = FDTD(Parameters(*shape), courant_dt(1, 1, 1), prec)
    for c, a in enumerate(random_state(int(np.prod(shape)), seed)):
        g.upload(c, a * (1.0 + c))
    g.set_point_sources([dict(i=1, j=2, k=1, comp=2, kind=0, amp=3.0)])
    g.update_fields(5)
    return g


def sums(fn, g):
    out = (ctypes.c_double * 6)()
    _lib.check(fn(g.handle, out))
    return list(out)


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32, Precision.FP16])
def test_noalloc_sum_squares_equals_allocating_entry_point(prec):
    L = _lib.load()
    big = grid((192, 160, 128), prec, 1)   # the reduction's full 1024 workgroups
    small = grid((24, 10, 6), prec, 2)     # 6 workgroups
    for g in (big, small, big):
        a = sums(L.fdtd_sum_squares_noalloc, g)
        b = sums(L.fdtd_sum_squares, g)
        print(prec.name, g.cells, a, flush=True)
        assert a == b                       # bit-identical
        assert all(x > 0 for x in a)
    # the state after more steps: still equal, and changed
    before = sums(L.fdtd_sum_squares_noalloc, small)
    small.update_fields(3)
    after = sums(L.fdtd_sum_squares_noalloc, small)
    assert after == sums(L.fdtd_sum_squares, small)
    assert after != before
    big.close()
    small.close()

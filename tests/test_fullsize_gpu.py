"""Size-independent properties at the benchmark's full sizes.

The oracle finishes small grids in seconds; at the benchmark sizes (512^3,
the headline workload; 1024x1024x256, the temporal-blocking workload) the
HIP path is checked through properties the Yee step has exactly, whatever
the size:

* translation equivariance: on a periodic grid with uniform coefficients,
  moving the point source by one cell moves every field by one cell,
  bitwise (FP32, and FP16 with in-run rescaling);
* linearity under a power-of-two scale: doubling the source doubles every
  field exactly in FP64 (all operations are linear, scaling by 2 is exact,
  and the fields stay far from the subnormal range);
* temporally blocked passes equal single sweeps bitwise, and FP32 single
  sweeps equal the three-sweep structure bitwise;
* a plane wave along x stays bitwise uniform in y and z, across strip
  tiles, wavefront boundaries, z chunks and the periodic wrap;
* the B update is minus a forward-difference curl of E, whose
  forward-difference divergence vanishes identically: from zero fields
  with the source in E only, div B stays at rounding level.

Planes are compared on the host (one plane is Ni*Nj values); whole grids
by checksum.
"""
import numpy as np
import pytest

from fdtd_method_amd import FDTD, Component, Parameters, Precision
from oracle.oracle import C

pytestmark = pytest.mark.gpu

DT = 0.5 / (C * np.sqrt(3.0))


def headline_source(i, j, k, amp=1.0):
    """The 512^3 workload's Gaussian-modulated Jz point source."""
    return dict(i=i, j=j, k=k, comp=2, kind=1, amp=amp, period=20 * DT, t0=120 * DT,
                sigma=40 * DT, phase=0.0)


# Modes seeded before stepping in the bitwise blocked-vs-single comparisons,
# so that every tile, wavefront boundary, z chunk and the periodic wrap
# carry non-zero data (a point source alone reaches a few dozen cells).
SEED_WAVES = [dict(m=(3, 1, 2), amp=0.5, phase=0.3, pol=(0.0, 1.0, 0.2)),
              dict(m=(-2, 5, 1), amp=0.3, phase=1.1, pol=(1.0, 0.0, 0.5)),
              dict(m=(1, -4, 7), amp=0.2, phase=2.0, pol=(0.3, 0.6, 1.0))]


def make(shape, src, prec=Precision.FP32, blocking=None, waves=False):
    g = FDTD(Parameters(*shape), DT, prec)
    if waves:
        if prec == Precision.FP16:
            g.set_scales([2.0 ** 10] * 6)
        for w in SEED_WAVES:
            g.add_plane_wave(w["m"], w["amp"], w["phase"], w["pol"])
    g.set_point_sources([src])
    if prec == Precision.FP16:
        # scales sized for the seeded modes, or (fields starting at zero) for
        # one step's injection, as the benchmark does, so that FP16 holds the
        # fields
        g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * DT * 1e-2)
    # blocking None: single sweeps (grids this large default to strip passes)
    g.set_time_blocking_strips(*(blocking or (0,)))
    return g


PLANES = (250, 256, 257, 300)
COMPS = (Component.EX, Component.EZ, Component.BY)


def test_translation_equivariance_512():
    n = 512
    steps = 130  # past the envelope's peak at 120 dt
    a = make((n, n, n), headline_source(256, 256, 256), blocking=(4, 1, 2, 16, 0))
    b = make((n, n, n), headline_source(257, 256, 256), blocking=(4, 1, 2, 16, 0))
    print("blocking", a.get_time_blocking_strips(), flush=True)
    a.update_fields(steps)
    b.update_fields(steps)
    for c in COMPS:
        for k in PLANES:
            pa = a.download_plane(c, k).reshape(n, n)  # [j, i]
            pb = b.download_plane(c, k).reshape(n, n)
            assert np.array_equal(np.roll(pa, 1, axis=1), pb), (c, k)
    assert np.abs(a.download_plane(Component.EZ, 256)).max() > 0
    a.close()
    b.close()


def test_linearity_512():
    # FP64: the fields stay far above the subnormal range, where rounding
    # would not commute with the doubling
    n = 512
    steps = 60
    a = make((n, n, n), headline_source(256, 256, 256, amp=1.0), Precision.FP64, (3, 1, 2, 8, 0))
    b = make((n, n, n), headline_source(256, 256, 256, amp=2.0), Precision.FP64, (3, 1, 2, 8, 0))
    a.update_fields(steps)
    b.update_fields(steps)
    for c in COMPS:
        for k in PLANES:
            pa = a.download_plane(c, k)
            pb = b.download_plane(c, k)
            assert np.array_equal(2.0 * pa, pb), (c, k)
    a.close()
    b.close()


@pytest.mark.parametrize("prec,blocking", [(Precision.FP32, (4, 1, 2, 16, 0)),
                                           (Precision.FP32, (3, 1, 4, 8, 0)),
                                           (Precision.FP64, (3, 1, 2, 8, 0)),
                                           (Precision.FP16, (4, 1, 2, 16, 0))])
def test_blocked_equals_single_sweeps_512(prec, blocking):
    n = 512
    src = headline_source(256, 256, 256)
    a = make((n, n, n), src, prec, waves=True)
    b = make((n, n, n), src, prec, blocking, waves=True)
    assert a.checksum() == b.checksum()
    L = blocking[0]
    for steps in (4 * L + 1, L):  # remainder sweeps, then whole passes with the carried step
        a.update_fields(steps)
        b.update_fields(steps)
        assert a.checksum() == b.checksum(), steps
    # every tile and z chunk carried data: no plane of any component is zero
    for c in (Component.EX, Component.EY, Component.EZ, Component.BX, Component.BY,
              Component.BZ):
        for k in (0, 63, 64, 200, 511):
            assert np.count_nonzero(b.download_plane(c, k)) > 0.99 * n * n, (c, k)
    a.close()
    b.close()


def test_blocked_equals_single_sweeps_1024x1024x256():
    shape = (1024, 1024, 256)
    src = headline_source(300, 700, 100)
    a = make(shape, src, waves=True)
    b = make(shape, src, blocking=(4, 1, 2, 16, 0), waves=True)
    a.update_fields(13)
    b.update_fields(13)
    assert a.checksum() == b.checksum()
    a.close()
    b.close()


def test_fused_equals_three_sweeps_fp32_512():
    n = 512
    src = headline_source(256, 256, 256)
    a = make((n, n, n), src)
    b = make((n, n, n), src)
    a.update_fields(7)
    b.update_fields_3sweep(7)
    assert a.checksum() == b.checksum()
    a.close()
    b.close()


@pytest.mark.parametrize("prec,blocking,tol", [(Precision.FP64, (4, 1, 2, 8, 0), 1e-12),
                                               (Precision.FP32, (4, 1, 2, 16, 0), 1e-5)])
def test_divergence_of_B_stays_zero_512(prec, blocking, tol):
    n = 512
    g = make((n, n, n), headline_source(256, 256, 256), prec, blocking)
    g.update_fields(130)  # the wave front is ~37 cells out
    for k in (240, 255, 256, 280):
        bx = g.download_plane(Component.BX, k)  # [j, i], d = 1
        by = g.download_plane(Component.BY, k)
        bz0 = g.download_plane(Component.BZ, k)
        bz1 = g.download_plane(Component.BZ, k + 1)
        div = (np.roll(bx, -1, axis=1) - bx) + (np.roll(by, -1, axis=0) - by) + (bz1 - bz0)
        scale = max(np.abs(bx).max(), np.abs(by).max(), np.abs(bz0).max())
        assert scale > 0, k
        assert np.abs(div).max() <= tol * scale, (k, np.abs(div).max() / scale)
    g.close()


@pytest.mark.parametrize("prec,blocking", [(Precision.FP32, (4, 1, 2, 16, 0)),
                                           (Precision.FP64, (4, 1, 2, 8, 0)),
                                           (Precision.FP32, None)])
def test_plane_wave_stays_transversely_uniform_512(prec, blocking):
    """A plane wave along x (E along y, B along z) depends on i only: every
    cell of a row (and every row of every plane) takes the same operations,
    so the fields stay bitwise uniform in j and k, and Ex, Ez, Bx, By stay
    exactly zero."""
    n = 512
    g = FDTD(Parameters(n, n, n), DT, prec)
    g.add_plane_wave((3, 0, 0), 1.0, 0.3, (0.0, 1.0, 0.0))
    g.set_time_blocking_strips(*(blocking or (0,)))
    g.update_fields(37)  # strip passes and a remainder sweep
    for c in (Component.EY, Component.BZ):
        first = None
        for k in (0, 63, 64, 255, 300, 511):
            p = g.download_plane(c, k)  # [j, i]
            assert (p == p[0:1, :]).all(), (c, k)
            if first is None:
                first = p[0]
                assert np.abs(first).max() > 0.5 * (1.0 if c == Component.EY else 0.0), c
            assert np.array_equal(p[0], first), (c, k)
    for c in (Component.EX, Component.EZ, Component.BX, Component.BY):
        for k in (0, 255, 511):
            assert not g.download_plane(c, k).any(), (c, k)
    g.close()


def test_translation_equivariance_fp16_rescaling_512():
    """FP16 storage with in-run rescaling every 20 steps through strip passes:
    the grid-wide maxima, hence every scale decision, are the same for a
    shifted source, so moving the source one cell moves every field one cell,
    bitwise."""
    n = 512
    steps = 130
    grids = []
    for i in (256, 257):
        g = make((n, n, n), headline_source(i, 256, 256), Precision.FP16, (4, 1, 2, 16, 0))
        g.set_autoscale_interval(20, 10)
        s0 = g.get_scales()
        g.update_fields(steps)
        assert g.get_scales() != s0  # the rescaling passes ran
        grids.append(g)
    a, b = grids
    assert a.get_scales() == b.get_scales()
    for c in COMPS:
        for k in PLANES:
            pa = a.download_plane(c, k).reshape(n, n)
            pb = b.download_plane(c, k).reshape(n, n)
            assert np.array_equal(np.roll(pa, 1, axis=1), pb), (c, k)
    assert np.abs(a.download_plane(Component.EZ, 256)).max() > 0
    a.close()
    b.close()


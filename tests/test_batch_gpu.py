"""Batched solves through the C-ABI (fdtd_batch_*): the persistent batched
sweep must give every grid exactly the fields of its own single-grid run
(and, in FP64, of the oracle), directly and through captured HIP graphs."""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Batch, Parameters, Precision, plane_wave_batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(32, 48, 16), (48, 32, 64), (16, 16, 16), (64, 20, 24), (160, 8, 4), (8, 3, 5)]


def make(shape, prec, seed, src=True):
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1, factor=0.5)
    g = FDTD(Parameters(*shape), dt, prec)
    for c, a in enumerate(random_state(n, seed)):
        g.upload(c, a)
    if src:
        g.set_point_sources([dict(i=shape[0] // 2, j=shape[1] // 3, k=shape[2] - 1, comp=2,
                                  kind=1, amp=1.0, period=20 * dt, t0=10 * dt, sigma=8 * dt)])
    return g


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
def test_batch_equals_single_grid_runs(prec):
    singles = [make(s, prec, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, prec, q) for q, s in enumerate(SHAPES)]
    b = Batch(batched)
    print("batch shape", b.shape, flush=True)
    for g in singles:
        g.update_fields(9)
    b.update_fields(9)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()
        assert g1.step == g2.step == 9


def test_batch_fp64_matches_oracle():
    shapes = SHAPES[:4]
    dt = courant_dt(1, 1, 1, factor=0.5)
    grids = [make(s, Precision.FP64, 10 + q, src=False) for q, s in enumerate(shapes)]
    b = Batch(grids)
    b.update_fields(6)
    for q, s in enumerate(shapes):
        r = O.OracleGrid(*s, 1.0, 1.0, 1.0, dt)
        for c, a in enumerate(random_state(int(np.prod(s)), 10 + q)):
            r.fields()[c][:] = a
        r.update_fields(6)
        out = grids[q].fields()
        for c in range(6):
            assert np.array_equal(out[c], r.fields()[c])


@pytest.mark.parametrize("nsteps,replays", [(8, 3), (5, 3)])
def test_batch_graph_replays(nsteps, replays):
    """A captured step sequence replayed; odd lengths swap the buffers
    between replays and the device step counter keeps source time going."""
    singles = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    b = Batch(batched)
    b.update_fields_graph(nsteps, replays)
    for g in singles:
        g.update_fields(nsteps * replays)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()
        assert g2.step == nsteps * replays
    # a second call continues in time
    b.update_fields_graph(nsteps, 1)
    for g in singles:
        g.update_fields(nsteps)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()


def test_batch_shapes_and_autotune_keep_results():
    singles = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    b = Batch(batched)
    tuned, ms = b.autotune(reps=2)
    assert ms > 0 and tuned["work_items"] > 0
    for shape in [(4, 256, 8192), (2, 512, 65536), (1, 1024, 32768)]:
        b.set_shape(*shape)
        b.update_fields(3)
    for g in singles:
        g.update_fields(9)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()


def test_plane_wave_batch_workload_small():
    grids, desc = plane_wave_batch(n_grids=12, seed=5)
    ref, _ = plane_wave_batch(n_grids=12, seed=5)
    b = Batch(grids)
    b.update_fields_graph(10, 2)
    for g in ref:
        g.update_fields(20)
    for g1, g2 in zip(ref, grids):
        assert g1.checksum() == g2.checksum()
    ms_b, ms_s = b.time(4)
    print("batched ms/step", ms_b, "sequential", ms_s)
    assert ms_b > 0 and ms_s > 0


@pytest.mark.parametrize("prec,strips", [(Precision.FP64, (2, 1, 2, 8)),
                                         (Precision.FP32, (4, 1, 2, 16)),
                                         (Precision.FP32, (3, 1, 4, 4)),
                                         (Precision.FP32, (4, 2, 2, 8)),
                                         (Precision.FP32, (1, 1, 2, 16)),
                                         (Precision.FP16, (3, 1, 2, 16))])
def test_batch_strips_equal_single_grid_runs(prec, strips):
    """Batched strip passes (L steps per persistent launch, plane-sorted
    sources per grid) plus remainder sweeps give every grid its single-grid
    fields bitwise."""
    singles = [make(s, prec, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, prec, q) for q, s in enumerate(SHAPES)]
    for g in singles:
        g.set_step_carry(False)  # a batch call starts from the B/2 prologue
    b = Batch(batched)
    b.set_time_blocking_strips(*strips)
    cfg = b.get_time_blocking_strips()
    print("strips", cfg, flush=True)
    assert (cfg["L"], cfg["V"], cfg["R"], cfg["W"]) == strips
    L = strips[0]
    for nsteps in (2 * L + 1, L):
        for g in singles:
            g.update_fields(nsteps)
        b.update_fields(nsteps)
        for g1, g2 in zip(singles, batched):
            assert g1.checksum() == g2.checksum(), nsteps


def test_batch_strips_whole_grid_chunks():
    """A batch whose whole-grid tiles just fill four rounds of the device's
    256 resident workgroups (73 grids x 14 tiles = 1022) runs each grid as
    one z chunk (Nk = 96, which a batch of a few grids cuts into two
    48-plane chunks); the first grids still get their single-grid fields
    bitwise."""
    shape, ngrids, nsrc = (64, 160, 96), 73, 3
    singles = [make(shape, Precision.FP32, q) for q in range(nsrc)]
    batched = [make(shape, Precision.FP32, q) for q in range(nsrc)]
    batched += [FDTD(Parameters(*shape), courant_dt(1, 1, 1, factor=0.5), Precision.FP32)
                for _ in range(ngrids - nsrc)]
    for g in singles:
        g.set_step_carry(False)
    b = Batch(batched)
    b.set_time_blocking_strips(4, 1, 2, 16)
    cfg = b.get_time_blocking_strips()
    print("strips", cfg, flush=True)
    # 56 stored columns and 24 stored rows per tile: 2 x 7 tiles per chunk
    assert cfg["work_items"] == ngrids * 2 * 7
    for g in singles:
        g.update_fields(9)
    b.update_fields(9)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()
    for g in batched[nsrc:]:
        assert g.checksum() == batched[-1].checksum()
    b.close()
    for g in singles + batched:
        g.close()


@pytest.mark.parametrize("nsteps,replays,L", [(8, 3, 4), (5, 3, 2)])
def test_batch_strips_graph_replays(nsteps, replays, L):
    """Captured strip passes replayed; an odd number of launches per replay
    (passes plus remainder sweeps) swaps the buffers between replays."""
    singles = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    b = Batch(batched)
    b.set_time_blocking_strips(L, 1, 2, 8)
    b.update_fields_graph(nsteps, replays)
    for g in singles:
        g.update_fields(nsteps * replays)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()
        assert g2.step == nsteps * replays
    b.set_time_blocking_strips(0)
    b.update_fields(3)
    for g in singles:
        g.update_fields(3)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()


def test_batch_strips_reject_invalid_shapes():
    grids = [make(s, Precision.FP64, q) for q, s in enumerate(SHAPES[:2])]
    b = Batch(grids)
    for args in [(2, 1, 4, 8), (5, 1, 2, 8), (2, 1, 2, 1), (2, 1, 3, 8)]:
        with pytest.raises((ValueError, RuntimeError)):
            b.set_time_blocking_strips(*args)
    assert b.get_time_blocking_strips()["L"] == 0


@pytest.mark.parametrize("strips", [None, (4, 1, 2, 8)])
def test_batch_graph_recaptures_after_source_change(strips):
    """A captured graph bakes in the batch's source count (value-table
    stride, evaluation size): changing one grid's point sources between two
    graph calls of the same length (staying under the source-buffer
    capacity) must re-capture, and every grid keeps its single-grid fields
    bitwise."""
    singles = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    batched = [make(s, Precision.FP32, q) for q, s in enumerate(SHAPES)]
    b = Batch(batched)
    if strips:
        b.set_time_blocking_strips(*strips)
        for g in singles:
            g.set_step_carry(False)
    b.update_fields_graph(8, 2)
    for g in singles:
        g.update_fields(16)
    dt = courant_dt(1, 1, 1, factor=0.5)
    extra = [dict(i=1, j=2, k=0, comp=0, kind=1, amp=0.5, period=12 * dt, t0=4 * dt, sigma=0.0),
             dict(i=3, j=0, k=1, comp=1, kind=0, amp=0.25)]
    for grids in (singles, batched):
        s0 = SHAPES[0]
        grids[0].set_point_sources([dict(i=s0[0] // 2, j=s0[1] // 3, k=s0[2] - 1, comp=2, kind=1,
                                         amp=1.0, period=20 * dt, t0=10 * dt, sigma=8 * dt)]
                                   + extra)
        grids[3].set_point_sources([])
    b.update_fields_graph(8, 2)
    for g in singles:
        g.update_fields(16)
    for q, (g1, g2) in enumerate(zip(singles, batched)):
        assert g1.checksum() == g2.checksum(), q
        assert g1.step == g2.step == 32


@pytest.mark.parametrize("ngrids", [3, 100])
def test_batch_strips_chunked_grids(ngrids):
    """Batched strip passes with more than one z chunk per grid: a few
    (64, 160, 96) grids are cut into z chunks so that their tiles fill the
    device's resident workgroups, and so are 100 of them (the schedule
    model prefers two 48-plane chunks there).  The first grids carry data
    and must equal their single-grid runs bitwise."""
    shape, nsrc = (64, 160, 96), 3
    singles = [make(shape, Precision.FP32, q) for q in range(nsrc)]
    batched = [make(shape, Precision.FP32, q) for q in range(nsrc)]
    batched += [FDTD(Parameters(*shape), courant_dt(1, 1, 1, factor=0.5), Precision.FP32)
                for _ in range(ngrids - nsrc)]
    for g in singles:
        g.set_step_carry(False)
    b = Batch(batched)
    b.set_time_blocking_strips(4, 1, 2, 16)
    cfg = b.get_time_blocking_strips()
    print("strips", cfg, flush=True)
    # 2 x 7 tiles per z chunk, at least two chunks per grid
    assert cfg["work_items"] % (ngrids * 2 * 7) == 0
    assert cfg["work_items"] >= ngrids * 2 * 7 * 2
    for g in singles:
        g.update_fields(11)
    b.update_fields(11)
    for g1, g2 in zip(singles, batched):
        assert g1.checksum() == g2.checksum()
    b.close()
    for g in singles + batched:
        g.close()


def test_batch_256_grid_workload_graph_strips():
    """The 256-grid seeded workload (dims in {32, 48, ..., 160}, seed 2024,
    one plane wave and one Gaussian-modulated Jz source per grid), 500
    steps as graph-captured batched strip passes (L = 4, 20-step graph
    replayed 25 times): the 8 largest and 8 smallest grids equal their own
    single-grid runs bitwise."""
    grids, desc = plane_wave_batch(n_grids=256, seed=2024)
    cells = [g.cells for g in grids]
    order = np.argsort(cells, kind="stable")
    picked = [int(q) for q in list(order[:8]) + list(order[-8:])]
    ref, _ = plane_wave_batch(n_grids=256, seed=2024)
    keep = [ref[q] for q in picked]
    for q, g in enumerate(ref):
        if q not in picked:
            g.close()
    b = Batch(grids)
    b.set_time_blocking_strips(4, 1, 2, 16)
    print("strips", b.get_time_blocking_strips(), "cells", sum(cells), flush=True)
    b.update_fields_graph(20, 25)
    for g in keep:
        g.set_step_carry(False)
        g.update_fields(500)
    for q, g1 in zip(picked, keep):
        assert g1.checksum() == grids[q].checksum(), (q, desc[q]["shape"])
        assert grids[q].step == 500
    b.close()
    for g in grids + keep:
        g.close()

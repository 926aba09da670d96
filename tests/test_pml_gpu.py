"""Parity of the PML layer on the GPU (FDTD_PML: src/FDTD/FDTD_PML.cpp) with
the CPU oracle, through the C-ABI.

FP64 is bit-identical to the oracle (which matches the reference's compiled
PML class bit for bit) for the fused and three-sweep paths: totals and all
twelve split fields, on ragged shapes, layer widths 0..N/2 and every launch
shape.  FP32 is bit-identical between its fused and three-sweep paths and
within 1e-5 relative L2 of FP64; FP16 within 5e-3.
"""
import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, FDTD_PML, Batch, Parameters, Precision, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def pml_gpu(shape, spacing, dt, pml, state=None, precision=Precision.FP64, tune=None):
    g = FDTD_PML(Parameters(*shape, dx=spacing[0], dy=spacing[1], dz=spacing[2]), dt, pml,
                 precision)
    if tune:
        g.set_tuning(*tune)
    if state is not None:
        for c in range(len(state)):
            g.upload(c, state[c])
    return g


def pml_oracle(shape, spacing, dt, pml, state, J=None):
    r = O.OraclePmlGrid(*shape, *spacing, dt, pml)
    for c in range(6):
        r.fields()[c][:] = state[c]
    if J is not None:
        for c in range(3):
            r.J[c][:] = J[c]
    return r


def assert_same(g, r, splits=True):
    for c in range(6):
        out = g.download(c)
        assert np.array_equal(out, r.fields()[c]), \
            f"component {c}: {np.count_nonzero(out != r.fields()[c])} cells differ"
    if splits:
        for q in range(12):
            ref = (r.ES + r.BS)[q]
            out = g.split(q)
            assert np.array_equal(out, ref), \
                f"split {FDTD_PML.SPLITS[q]}: {np.count_nonzero(out != ref)} cells differ"


def rel_l2(out, ref):
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(b ** 2)) for b in ref)
    return np.sqrt(num / den)


CASES = [((16, 12, 10), 0.2), ((13, 9, 11), 0.3), ((10, 6, 4), 0.5), ((24, 20, 16), 0.25),
         ((7, 5, 3), 0.15), ((64, 48, 40), 0.2), ((33, 17, 9), 0.2), ((5, 12, 1), 0.25),
         ((40, 3, 30), 0.1)]


def spread_current(n, seed):
    J = np.zeros(n)
    J[:: max(1, n // 9)] = np.random.default_rng(seed).standard_normal(J[:: max(1, n // 9)].size)
    return J


@pytest.mark.parametrize("path", ["phases", "fused", "3sweep"])
@pytest.mark.parametrize("shape,pml", CASES, ids=lambda v: "x".join(map(str, v))
                         if isinstance(v, tuple) else str(v))
def test_pml_fp64_bit_exact(shape, pml, path):
    """Random fields and a dense J = Jx = Jy = Jz on cells of the main box and
    of the layer; 5 steps; totals and splits bit-exact.  fused: the fused
    PML sweep (the default); phases: the two in-place phase sweeps per step;
    3sweep: the reference's structure."""
    n = int(np.prod(shape))
    spacing = (1.0, 1.3, 0.7)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=n + 3)
    J = spread_current(n, seed=n)
    print(f"case shape={shape} pml={pml} path={path}", flush=True)
    g = pml_gpu(shape, spacing, dt, pml, st + [J, J, J])
    r = pml_oracle(shape, spacing, dt, pml, st, [J, J, J])
    g.set_pml_path(1 if path == "phases" else 0)
    if path == "3sweep":
        g.update_fields_3sweep(5)
    else:
        g.update_fields(5)
    r.update_fields(5)
    assert g.pml_widths == O.pml_sizes(*shape, pml)
    assert_same(g, r)


LAUNCH = [(1, 1, 1), (1, 3, 4), (2, 2, 5), (2, 7, 18), (1, 17, 3), (1, 11, 7), (2, 23, 2)]


@pytest.mark.parametrize("tune", LAUNCH, ids=lambda t: "V%d_TY%d_k%d" % t)
def test_pml_fp64_launch_shapes(tune):
    """Tile heights, z chunks and lane widths put halo rows and chunk-boundary
    planes inside and across the layer."""
    shape, pml, spacing = (40, 24, 18), 0.2, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    st = random_state(n, seed=77)
    print(f"case tune={tune}", flush=True)
    g = pml_gpu(shape, spacing, dt, pml, st, tune=tune)
    g.set_pml_path(0)  # the tuning shapes the fused sweep
    r = pml_oracle(shape, spacing, dt, pml, st)
    for steps in (1, 4):
        g.update_fields(steps)
        r.update_fields(steps)
        assert_same(g, r)


PHASE_SHAPES = [(1, 1, 1), (1, 3, 4), (2, 2, 5), (2, 7, 18), (2, 1, 3), (1, 5, 2), (1, 23, 7),
                (2, 24, 18)]


@pytest.mark.parametrize("shape_v", PHASE_SHAPES, ids=lambda t: "V%d_TY%d_k%d" % t)
def test_pml_fp64_phase_shapes(shape_v):
    """The phase sweeps' tile heights (tiles of one row, of a ragged last
    tile, of the whole grid), z chunks (one plane to whole) and lane widths
    (FP64 packs of 1 and 2 cells; 4-cell packs run in the FP32 and FP16
    tests against the fused sweep) put tile boundaries and chunk boundaries
    inside and across the layer."""
    shape, pml, spacing = (40, 24, 18), 0.2, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    st = random_state(n, seed=78)
    print(f"case phases V,TY,kchunk={shape_v}", flush=True)
    g = pml_gpu(shape, spacing, dt, pml, st)
    g.set_pml_path(1, *shape_v)
    assert g.get_pml_path() == dict(path=1, V=shape_v[0], TY=shape_v[1], kchunk=shape_v[2])
    r = pml_oracle(shape, spacing, dt, pml, st)
    for steps in (1, 4):
        g.update_fields(steps)
        r.update_fields(steps)
        assert_same(g, r)


@pytest.mark.parametrize("prec", [Precision.FP32, Precision.FP16])
def test_pml_phases_equal_fused_sweep(prec):
    """The two phase sweeps per step and the fused PML sweep give the same
    bits in FP32, with point sources in the main box and the layer.  With
    FP16 storage (power-of-two scales, then in-run rescaling, whose
    rescaling sweeps are fused ones in both runs) the phase path's B update
    reads E_{n+1} as stored, rounded to FP16, where the fused sweep uses it
    before rounding: the two agree to FP16 rounding (relative L2 <= 5e-3,
    3x the measured 1.7e-3), with the same scale decisions."""
    shape, pml, spacing = (48, 40, 36), 0.2, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = 0.5 / (O.C * np.sqrt(3.0))
    st = [0.01 * x for x in random_state(n, seed=31)]
    src = [dict(i=24, j=20, k=18, comp=2, kind=1, amp=1.0, period=20 * dt, t0=10 * dt,
                sigma=15 * dt),
           dict(i=2, j=3, k=4, comp=0, kind=0, amp=0.5)]
    runs = []
    for path in (1, 0):
        g = pml_gpu(shape, spacing, dt, pml, st, precision=prec)
        g.set_point_sources(src)
        if prec == Precision.FP16:
            g.autoscale(target_log2=10)
        g.set_pml_path(path)
        g.update_fields(7)
        g.update_fields(1)
        if prec == Precision.FP16:
            g.set_autoscale_interval(4, 10)
        g.update_fields(13)
        runs.append(g)
    a, b = runs
    assert a.get_scales() == b.get_scales()
    if prec == Precision.FP16:
        err = rel_l2([a.download(c) for c in range(6)], [b.download(c) for c in range(6)])
        print("fp16 phases vs fused rel L2", err, flush=True)
        assert 0 < err <= 5e-3
        return
    assert a.checksum() == b.checksum()
    for q in range(12):
        assert np.array_equal(a.split(q), b.split(q)), FDTD_PML.SPLITS[q]


def test_pml_fp64_matches_golden(golden_dir):
    """The reference class's own dump (tests/golden/ref_pml_16x12x10_random_40steps.npz)."""
    import os
    f = np.load(os.path.join(golden_dir, "ref_pml_16x12x10_random_40steps.npz"))
    shape = tuple(int(x) for x in f["dims"])
    J = f["J_in"]
    st = list(f["E_in"]) + list(f["B_in"])
    g = pml_gpu(shape, tuple(f["spacing"]), float(f["dt"]), float(f["pml_percent"]),
                st + [J, J, J])
    g.update_fields(int(f["nsteps"]))
    for c in range(3):
        assert np.array_equal(g.download(c), f["E_out"][c])
        assert np.array_equal(g.download(3 + c), f["B_out"][c])


def test_pml_point_sources_and_sparse_currents():
    """Waveform sources evaluated on the device (FP64 sin/exp: 1e-12 relative
    L2) and sparse explicit currents (bit-exact), in the main box and the
    layer (where they have no effect)."""
    shape, pml, spacing = (32, 28, 24), 0.2, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    Ni, Nj = shape[0], shape[1]
    src = [dict(i=16, j=14, k=12, comp=2, kind=1, amp=1.0, period=20 * dt, t0=30 * dt,
                sigma=10 * dt),
           dict(i=1, j=2, k=3, comp=0, kind=1, amp=2.0, period=15 * dt, t0=0.0, sigma=0.0)]
    g = pml_gpu(shape, spacing, dt, pml)
    g.set_point_sources(src)
    r = O.OraclePmlGrid(*shape, *spacing, dt, pml)
    g.update_fields(60)
    r.run_point_sources(src, 60)
    assert rel_l2([g.download(c) for c in range(6)], r.fields()[:6]) < 1e-12
    # sparse values: kind-0 currents held constant, bit-exact
    g2 = pml_gpu(shape, spacing, dt, pml)
    cells = [15 + 13 * Ni + 11 * Ni * Nj, 0]
    g2.set_current_values(cells, [7, 7], [0.75, 5.0])  # JY
    J = np.zeros(n)
    J[cells] = [0.75, 5.0]
    r2 = pml_oracle(shape, spacing, dt, pml, [np.zeros(n)] * 6, [np.zeros(n), J, np.zeros(n)])
    g2.update_fields(9)
    r2.update_fields(9)
    assert_same(g2, r2)


def test_pml_multi_step_and_host_writes():
    """update_fields(n) == n x update_fields(); totals written into layer cells
    between steps behave as in the reference (splits are kept)."""
    shape, pml, spacing = (20, 18, 14), 0.25, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    st = random_state(n, seed=4)
    a = pml_gpu(shape, spacing, dt, pml, st)
    b = pml_gpu(shape, spacing, dt, pml, st)
    r = pml_oracle(shape, spacing, dt, pml, st)
    for _ in range(4):
        a.update_fields(1)
    b.update_fields(4)
    r.update_fields(4)
    assert_same(a, r)
    assert_same(b, r)
    ex = a.download(0)
    ex[:50] = 0.5
    a.upload(0, ex)
    r.E[0][:50] = 0.5
    a.update_fields(3)
    r.update_fields(3)
    assert_same(a, r)


def test_pml_zero_width_is_the_periodic_grid():
    shape, spacing = (16, 16, 16), (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    st = random_state(n, seed=8)
    for pml in (0.0, 0.05):
        g = pml_gpu(shape, spacing, dt, pml, st)
        h = FDTD(Parameters(*shape), dt)
        for c in range(6):
            h.upload(c, st[c])
        g.update_fields(7)
        h.update_fields(7)
        assert g.pml_widths == (0, 0, 0)
        assert g.checksum() == h.checksum()
        assert not g.split("Exy").any()


def test_pml_fp32_fused_equals_3sweep_and_tracks_fp64():
    """A pulse absorbed by the layer: FP32 fused == FP32 three-sweep bitwise;
    the FP32 error is held to 1e-5 of the FP64 field norm while the pulse is
    inside the grid (120 steps), and after absorption (300 steps) to 1e-5 of
    that same norm (the residual field is orders of magnitude smaller, so
    an error relative to it would measure FP32 rounding of the absorbed
    pulse, not the solver)."""
    shape, pml, spacing = (48, 40, 36), 0.2, (1.0, 1.0, 1.0)
    dt = 0.5 / (O.C * np.sqrt(3.0))
    src = [dict(i=24, j=20, k=18, comp=2, kind=1, amp=1.0, period=20 * dt, t0=60 * dt,
                sigma=15 * dt)]
    grids, out = {}, {}
    for name, prec in (("f64", Precision.FP64), ("f32", Precision.FP32), ("f32s", Precision.FP32)):
        g = grids[name] = pml_gpu(shape, spacing, dt, pml, precision=prec)
        g.set_point_sources(src)
    for steps in (120, 180):
        for name, g in grids.items():
            if name == "f32s":
                g.update_fields_3sweep(steps)
            else:
                g.update_fields(steps)
            out[name, steps] = [g.download(c) for c in range(6)]
        for c in range(6):
            assert np.array_equal(out["f32", steps][c], out["f32s", steps][c])
    norm = lambda f: np.sqrt(sum(float(np.sum(a * a)) for a in f))  # noqa: E731
    peak = norm(out["f64", 120])
    err120 = norm([a - b for a, b in zip(out["f32", 120], out["f64", 120])]) / peak
    err300 = norm([a - b for a, b in zip(out["f32", 180], out["f64", 180])]) / peak
    print("fp32 error / fp64 norm at 120 steps", err120, "at 300 steps", err300,
          "residual / peak", norm(out["f64", 180]) / peak)
    assert err120 < 1e-5
    assert err300 < 1e-5


def test_pml_fp16_with_scales_tracks_fp64():
    """FP16 storage with per-component power-of-two scales (autoscaled in the
    run; split fields share their totals' scales): error within 5e-3 of the
    FP64 field norm while the pulse is in the grid, and the layer still
    absorbs."""
    shape, pml, spacing = (48, 40, 36), 0.2, (1.0, 1.0, 1.0)
    dt = 0.5 / (O.C * np.sqrt(3.0))
    src = [dict(i=24, j=20, k=18, comp=2, kind=1, amp=1.0, period=20 * dt, t0=60 * dt,
                sigma=15 * dt)]
    ref = pml_gpu(shape, spacing, dt, pml)
    ref.set_point_sources(src)
    ref.update_fields(120)
    h = pml_gpu(shape, spacing, dt, pml, precision=Precision.FP16)
    h.set_point_sources(src)
    h.set_scales([2.0 ** 30] * 6)
    h.set_autoscale_interval(10, 10)
    h.update_fields(120)
    err = rel_l2([h.download(c) for c in range(6)], [ref.download(c) for c in range(6)])
    e0 = sum(h.sum_squares())
    h.update_fields(300)
    e1 = sum(h.sum_squares())
    print("fp16 vs fp64 rel L2", err, "energy", e0, e1, h.get_scales())
    assert err < 5e-3
    assert e1 < 1e-2 * e0


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_pml_absorbs_on_gpu(precision):
    N, pml = 48, 0.2
    dt = 0.5 / (O.C * np.sqrt(3.0))
    src = [dict(i=24, j=24, k=24, comp=2, kind=1, amp=1.0, period=20 * dt, t0=60 * dt,
                sigma=15 * dt)]
    g = pml_gpu((N, N, N), (1.0, 1.0, 1.0), dt, pml, precision=precision)
    g.set_point_sources(src)
    g.update_fields(150)
    e0 = sum(g.sum_squares())
    g.update_fields(350)
    e1 = sum(g.sum_squares())
    print("energy", e0, e1)
    assert e1 < 1e-3 * e0


def test_pml_autotune_keeps_state_and_restrictions():
    shape, pml, spacing = (64, 40, 30), 0.2, (1.0, 1.0, 1.0)
    n = int(np.prod(shape))
    dt = courant_dt(*spacing)
    st = random_state(n, seed=21)
    g = pml_gpu(shape, spacing, dt, pml, st)
    r = pml_oracle(shape, spacing, dt, pml, st)
    g.update_fields(3)
    r.update_fields(3)
    g.autotune(reps=2)
    g.update_fields(4)
    r.update_fields(4)
    assert_same(g, r)
    with pytest.raises(_lib.FDTDError):
        g.set_time_blocking(2)
    with pytest.raises(_lib.FDTDError):
        Batch([g])

"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

FP64 is required to be bit-identical to the oracle (and therefore to the
reference's compiled class, which the oracle matches bit for bit) for the
fused and the three-sweep paths, on ragged shapes and every launch shape.
Sources evaluated on the device (sin/exp in FP64) are held to 1e-12
relative L2.  FP32/FP16 are held to stated tolerances against FP64.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import courant_dt, random_state
from fdtd_method_amd import FDTD, Component, Parameters, Precision, _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def gpu_grid(shape, spacing, dt, state=None, precision=Precision.FP64, tune=None):
    g = FDTD(Parameters(*shape, dx=spacing[0], dy=spacing[1], dz=spacing[2]), dt, precision)
    if tune:
        g.set_tuning(*tune)
    if state is not None:
        for c in range(len(state)):
            g.upload(c, state[c])
    return g


def oracle_grid(shape, spacing, dt, state, J=None):
    r = O.OracleGrid(*shape, *spacing, dt)
    for c in range(6):
        r.fields()[c][:] = state[c]
    if J is not None:
        for c in range(3):
            r.J[c][:] = J[c]
    return r


def rel_l2(out, ref):
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(out, ref))
    den = sum(float(np.sum(b ** 2)) for b in ref)
    return np.sqrt(num / den)


SHAPES = [(16, 12, 10), (8, 6, 5), (64, 64, 64), (32, 48, 20), (4, 8, 16), (2, 3, 1),
          (1, 1, 7), (48, 5, 9), (160, 3, 2)]


@pytest.mark.parametrize("path", ["fused", "3sweep"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fp64_bit_exact(shape, path):
    n = int(np.prod(shape))
    spacing = (1.0, 1.3, 0.9)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=n)
    print(f"case shape={shape} path={path}", flush=True)
    g = gpu_grid(shape, spacing, dt, st)
    (g.update_fields if path == "fused" else g.update_fields_3sweep)(7)
    r = oracle_grid(shape, spacing, dt, st)
    r.update_fields(7)
    out = g.fields()
    for c in range(6):
        assert np.array_equal(out[c], r.fields()[c]), f"component {c}"


@pytest.mark.parametrize("tune", [(1, 1, 1), (1, 3, 5), (2, 1, 64), (2, 3, 7), (4, 2, 9),
                                  (2, 7, 3), (1, 15, 2)])
def test_fp64_bit_exact_all_launch_shapes(tune):
    shape = (64, 40, 24)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    st = random_state(n, seed=3)
    print(f"case tune={tune}", flush=True)
    g = gpu_grid(shape, spacing, dt, st, tune=tune)
    assert g.get_tuning() == tune
    g.update_fields(5)
    r = oracle_grid(shape, spacing, dt, st)
    r.update_fields(5)
    out = g.fields()
    for c in range(6):
        assert np.array_equal(out[c], r.fields()[c]), f"component {c}"


def test_golden_random_dump(golden_dir):
    f = np.load(os.path.join(golden_dir, "ref_16x12x10_random_50steps.npz"))
    shape = tuple(int(x) for x in f["dims"])
    g = gpu_grid(shape, tuple(f["spacing"]), float(f["dt"]), list(f["E_in"]) + list(f["B_in"]))
    for c in range(3):
        g.upload(6 + c, f["J_in"])
    g.update_fields(int(f["nsteps"]))
    out = g.fields()
    for c in range(3):
        assert np.array_equal(out[c], f["E_out"][c])
        assert np.array_equal(out[3 + c], f["B_out"][c])


def config1_state(N=64):
    n = N ** 3
    i = np.arange(n) % N
    ey = np.sin(2.0 * np.pi * i / N)
    bz = np.sin(2.0 * np.pi * (i + 0.5) / N) / O.C
    z = np.zeros(n)
    return [z.copy(), ey, z.copy(), z.copy(), z.copy(), bz]


def test_config1_plane_wave_matches_reference(golden_dir):
    """64^3 plane wave, 200 steps: all six FP64 fields bit-identical to the
    reference class's (SHA-256 of its dump)."""
    rec = json.load(open(os.path.join(golden_dir, "ref_config1_64cube_200steps.json")))
    st = config1_state(rec["dims"][0])
    g = gpu_grid(tuple(rec["dims"]), (rec["d"],) * 3, rec["dt"], st)
    g.update_fields(rec["nsteps"])
    assert [sha(a) for a in g.fields()] == rec["sha256"]


def test_config1_fp32_error_vs_fp64(golden_dir):
    rec = json.load(open(os.path.join(golden_dir, "ref_config1_64cube_200steps.json")))
    st = config1_state(rec["dims"][0])
    g64 = gpu_grid(tuple(rec["dims"]), (rec["d"],) * 3, rec["dt"], st)
    g32 = gpu_grid(tuple(rec["dims"]), (rec["d"],) * 3, rec["dt"], st, Precision.FP32)
    g64.update_fields(rec["nsteps"])
    g32.update_fields(rec["nsteps"])
    err = g32.compare(g64)
    rel = np.sqrt(sum(e[0] for e in err) / sum(e[1] for e in err))
    print("config1 fp32 relL2", rel)
    assert rel <= 1e-5


def test_sample_harness_sparse_currents(golden_dir):
    """The drivers' per-step box source written through get_field(J*)[idx]
    (sparse current values), then zeroed_currents(): bit-identical to the
    reference class (perf-tests/sample/sample.cpp:14-91)."""
    rec = json.load(open(os.path.join(golden_dir, "ref_sample_32_100.json")))
    n, it = rec["n"], rec["iterations"]
    p = O.sample_box(n)
    d = p["d"]
    g = FDTD(Parameters(n, n, n, -p["boundary"], p["boundary"], -p["boundary"], p["boundary"],
                        -p["boundary"], p["boundary"], d, d, d), p["dt"])
    cur_time = min(p["iterations"], it)
    for t in range(cur_time):
        for k in range(p["start"], p["stop"]):
            for j in range(p["start"], p["stop"]):
                for i in range(p["start"], p["stop"]):
                    idx = i + j * n + k * n * n
                    val = O.sample_current(i * d, j * d, k * d, (t + 1) * p["dt"], p["T"],
                                           p["Tx"], p["Ty"], p["Tz"])
                    for comp in (Component.JX, Component.JY, Component.JZ):
                        g.get_field(comp)[idx] = val
        g.update_fields()
    g.zeroed_currents()
    g.update_fields(it - cur_time)
    assert [sha(a) for a in g.fields()] == rec["sha256"]


def test_explicit_point_sources_bit_exact():
    shape = (24, 20, 16)
    n = int(np.prod(shape))
    spacing = (1.0, 1.25, 0.75)
    dt = courant_dt(*spacing)
    st = random_state(n, 21)
    src = [dict(i=5, j=7, k=3, comp=2, kind=0, amp=0.75), dict(i=0, j=19, k=15, comp=0, kind=0, amp=-2.0),
           dict(i=23, j=0, k=0, comp=1, kind=0, amp=1.5)]
    r = oracle_grid(shape, spacing, dt, st)
    r.run_point_sources(src, 9)
    g = gpu_grid(shape, spacing, dt, st)
    g.set_point_sources(src)
    g.update_fields(9)
    out = g.fields()
    for c in range(6):
        assert np.array_equal(out[c], r.fields()[c])


@pytest.mark.parametrize("path", ["fused", "3sweep"])
def test_waveform_sources_fp64(path):
    """Gaussian-modulated sinusoid evaluated on the device in FP64 at
    t=(n+1)dt: within 1e-12 relative L2 of the oracle (libm sin/exp)."""
    shape = (32, 24, 20)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing, factor=0.5)
    src = [dict(i=16, j=12, k=10, comp=2, kind=1, amp=1.0, period=20 * dt, t0=30 * dt, sigma=10 * dt),
           dict(i=3, j=4, k=5, comp=1, kind=1, amp=0.5, period=16 * dt, t0=0.0, sigma=0.0, phase=0.3)]
    zero = [np.zeros(n) for _ in range(6)]
    r = oracle_grid(shape, spacing, dt, zero)
    r.run_point_sources(src, 60)
    g = gpu_grid(shape, spacing, dt)
    g.set_point_sources(src)
    (g.update_fields if path == "fused" else g.update_fields_3sweep)(60)
    out = g.fields()
    rel = rel_l2(out, r.E + r.B)
    print("waveform relL2", rel, "exact", all(np.array_equal(a, b) for a, b in zip(out, r.E + r.B)))
    assert rel <= 1e-12


def test_many_sources_per_plane_table():
    """40 sources (kinds 0 and 1) spread over planes 0..Nk-1, several per
    plane, two sharing a cell and component: the per-plane source ranges of
    the fused sweep give the oracle's result (1e-12 relative L2; coincident
    sources add cur*J per source, the oracle sums J first), the three-sweep
    path's result bitwise, and a batch run's result bitwise."""
    from fdtd_method_amd import Batch
    shape = (28, 20, 18)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1, factor=0.5)
    rng = np.random.default_rng(77)
    src = []
    for q in range(38):
        k = int(rng.choice([0, shape[2] - 1, 5, 5, 9])) if q < 12 else int(rng.integers(shape[2]))
        src.append(dict(i=int(rng.integers(shape[0])), j=int(rng.integers(shape[1])), k=k,
                        comp=int(rng.integers(3)), kind=int(q % 2), amp=float(rng.uniform(-1, 1)),
                        period=float(rng.uniform(10, 30)) * dt, t0=20 * dt, sigma=8 * dt))
    src += [dict(i=4, j=6, k=7, comp=2, kind=0, amp=0.3), dict(i=4, j=6, k=7, comp=2, kind=0, amp=-1.1)]
    zero = [np.zeros(n) for _ in range(6)]
    r = oracle_grid(shape, (1, 1, 1), dt, zero)
    r.run_point_sources(src, 30)
    grids = []
    for path in ("fused", "3sweep", "batch"):
        g = gpu_grid(shape, (1, 1, 1), dt)
        g.set_point_sources(src)
        if path == "fused":
            g.update_fields(30)
        elif path == "3sweep":
            g.update_fields_3sweep(30)
        else:
            Batch([g]).update_fields(30)
        grids.append(g.fields())
    assert rel_l2(grids[0], r.E + r.B) <= 1e-12
    for c in range(6):
        assert np.array_equal(grids[0][c], grids[1][c]), f"3-sweep, component {c}"
        assert np.array_equal(grids[0][c], grids[2][c]), f"batch, component {c}"


def test_dense_and_sparse_currents_bit_exact():
    """Full J uploads (dense path) and more sparse cells than the sparse list
    holds (switch to dense) give the oracle's persistent-J result."""
    shape = (20, 18, 12)
    n = int(np.prod(shape))
    spacing = (1.0, 1.0, 1.0)
    dt = courant_dt(*spacing)
    st = random_state(n, 4, count=9)
    r = oracle_grid(shape, spacing, dt, st[:6], st[6:])
    r.update_fields(6)
    g = gpu_grid(shape, spacing, dt, st[:6])
    for c in range(3):
        g.upload(6 + c, st[6 + c])
    g.update_fields(6)
    out = g.fields()
    for c in range(6):
        assert np.array_equal(out[c], r.fields()[c])
    # sparse -> dense switch
    rng = np.random.default_rng(8)
    cells = rng.choice(n, 150, replace=False)
    J = [np.zeros(n) for _ in range(3)]
    vals = rng.standard_normal(150)
    comps = rng.integers(0, 3, 150)
    for q in range(150):
        J[comps[q]][cells[q]] = vals[q]
    r2 = oracle_grid(shape, spacing, dt, st[:6], J)
    r2.update_fields(4)
    g2 = gpu_grid(shape, spacing, dt, st[:6])
    g2.set_current_values(cells, comps + 6, vals)
    g2.update_fields(4)
    out = g2.fields()
    for c in range(6):
        assert np.array_equal(out[c], r2.fields()[c])
    for c in range(3):
        assert np.array_equal(g2.download(6 + c), J[c])
    # writes into dense J: one call with repeated cells (the last value
    # wins, as in element-by-element assignment), in FP64 and FP32
    for prec in (Precision.FP64, Precision.FP32):
        g3 = gpu_grid(shape, spacing, dt, st[:6], prec)
        g3.set_current_values(cells, comps + 6, vals)
        rep = np.concatenate([cells[:40], cells[:40], cells[100:120]])
        rcomp = np.concatenate([comps[:40], comps[:40], comps[100:120]])
        rval = rng.standard_normal(rep.size)
        g3.set_current_values(rep, rcomp + 6, rval)
        J3 = [a.copy() for a in J]
        for q in range(rep.size):
            J3[rcomp[q]][rep[q]] = rval[q]
        for c in range(3):
            want = J3[c] if prec == Precision.FP64 else J3[c].astype(np.float32).astype(np.float64)
            assert np.array_equal(g3.download(6 + c), want), (prec, c)
        if prec == Precision.FP64:
            r3 = oracle_grid(shape, spacing, dt, st[:6], J3)
            r3.update_fields(3)
            g3.update_fields(3)
            out = g3.fields()
            for c in range(6):
                assert np.array_equal(out[c], r3.fields()[c])


def test_currents_persist_and_zero():
    shape = (10, 8, 6)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    g = gpu_grid(shape, (1, 1, 1), dt)
    g.get_field(Component.JZ)[17] = 0.5
    g.update_fields(1)
    cur = -4.0 * O.PI * dt
    assert g.get_field(Component.EZ)[17] == cur * 0.5
    assert g.get_field(Component.JZ)[17] == 0.5
    g.zeroed_currents()
    assert not np.any(g.download(Component.JZ))
    r = oracle_grid(shape, (1, 1, 1), dt, [np.zeros(n)] * 6)
    r.J[2][17] = 0.5
    r.update_fields(1)
    r.J[2][17] = 0.0
    r.update_fields(3)
    g.update_fields(3)
    out = g.fields()
    for c in range(6):
        assert np.array_equal(out[c], r.fields()[c])


def test_multi_step_equals_repeated_single_steps():
    shape = (40, 20, 10)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 12)
    a = gpu_grid(shape, (1, 1, 1), dt, st)
    b = gpu_grid(shape, (1, 1, 1), dt, st)
    for _ in range(6):
        a.update_fields(1)
    b.update_fields(6)
    assert a.checksum() == b.checksum()
    assert a.step == b.step == 6


def test_fp32_fused_equals_fp32_three_sweep():
    """Without FMA the fused (B+d)+d and the three in-place sweeps round the
    same way in FP32 too."""
    shape = (48, 40, 30)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 13)
    a = gpu_grid(shape, (1, 1, 1), dt, st, Precision.FP32)
    b = gpu_grid(shape, (1, 1, 1), dt, st, Precision.FP32)
    a.update_fields(11)
    b.update_fields_3sweep(11)
    assert a.checksum() == b.checksum()


def test_fp16_plane_wave_error():
    """FP16 storage (FP32 compute) on a smooth plane wave: relative L2 error
    against FP64 after 100 steps stays below 5e-3 (measured 1.5e-3)."""
    N = 48
    n = N ** 3
    st = config1_state(N)
    st[5] = st[5] * O.C  # Gaussian units: |B| = |E| for a plane wave
    dt = courant_dt(1, 1, 1, factor=0.5)
    g64 = gpu_grid((N, N, N), (1, 1, 1), dt, st)
    g16 = gpu_grid((N, N, N), (1, 1, 1), dt, st, Precision.FP16)
    g64.update_fields(100)
    g16.update_fields(100)
    err = g16.compare(g64)
    rel = np.sqrt(sum(e[0] for e in err) / sum(e[1] for e in err))
    print("fp16 plane wave relL2", rel)
    assert rel <= 5e-3


def test_fp16_scales_are_exact_and_restored():
    shape = (16, 8, 8)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = [a * 1e-9 for a in random_state(n, 2)]
    g = gpu_grid(shape, (1, 1, 1), dt, None, Precision.FP16)
    g.set_scales([2.0 ** 30] * 6)
    for c in range(6):
        g.upload(c, st[c])
    before = g.fields()
    # upward power-of-two rescaling is exact even for FP16 subnormals
    g.set_scales([2.0 ** 31, 2.0 ** 32, 2.0 ** 30, 2.0 ** 31, 2.0 ** 30, 2.0 ** 32])
    after = g.fields()
    for c in range(6):
        assert np.array_equal(before[c], after[c])
    g.autoscale(target_log2=12)
    sc = g.get_scales()
    mx = g.max_abs()
    for c in range(6):
        assert 2.0 ** 11 <= mx[c] * sc[c] <= 2.0 ** 12
    with pytest.raises(_lib.FDTDError):
        g.set_scales([3.0] * 6)


def test_fp16_in_run_rescaling():
    """FP16 with rescaling every 5 sweeps, inside the run (the sweep before a
    decision point tracks max |stored|, the next one stores its output
    multiplied by new/old power-of-two factors): a growing point-source field
    keeps every component's stored max in range (scales move with it) and the
    result stays within 5e-3 of FP64, about three times the measured error."""
    N = 40
    dt = courant_dt(1, 1, 1, factor=0.5)
    src = [dict(i=20, j=20, k=20, comp=2, kind=1, amp=1.0, period=20 * dt, t0=60 * dt,
                sigma=20 * dt)]
    ref = gpu_grid((N, N, N), (1, 1, 1), dt)
    ref.set_point_sources(src)
    ref.update_fields(90)
    g = gpu_grid((N, N, N), (1, 1, 1), dt, None, Precision.FP16)
    g.set_point_sources(src)
    g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * dt * 1e-2)
    s0 = g.get_scales()
    g.set_autoscale_interval(5, 10)
    g.update_fields(90)
    s1 = g.get_scales()
    assert s1 != s0
    mx = g.max_abs()
    # rescaled to max in [2^9, 2^10) at the last decision point (step 85);
    # the field grows ~2x per 5 steps here, far from FP16's 2^16 limit
    for c in (0, 1, 2, 3, 4):   # Bz stays ~0 under a Jz source
        assert mx[c] * s1[c] <= 2.0 ** 14, (c, mx[c] * s1[c])
        assert mx[c] * s1[c] >= 2.0 ** 6, (c, mx[c] * s1[c])
    err = g.compare(ref)
    rel = np.sqrt(sum(e[0] for e in err) / sum(e[1] for e in err))
    print("fp16 in-run rescaling relL2", rel, s0, s1)
    assert rel <= 5e-3


def test_diagnostics_match_oracle():
    shape = (30, 20, 10)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1)
    st = random_state(n, 14)
    g = gpu_grid(shape, (1, 1, 1), dt, st)
    r = oracle_grid(shape, (1, 1, 1), dt, st)
    g.update_fields(3)
    r.update_fields(3)
    np.testing.assert_allclose(g.sum_squares(), r.sum_squares(), rtol=1e-12)
    np.testing.assert_array_equal(g.max_abs(), [np.max(np.abs(a)) for a in r.E + r.B])
    g32 = gpu_grid(shape, (1, 1, 1), dt, st, Precision.FP32)
    g32.update_fields(3)
    err = g32.compare(g)
    f32 = g32.fields()
    for c in range(6):
        d = f32[c] - r.fields()[c]
        np.testing.assert_allclose(err[c][0], np.sum(d * d), rtol=1e-10)
        np.testing.assert_allclose(err[c][1], np.sum(r.fields()[c] ** 2), rtol=1e-12)
        assert err[c][2] == np.max(np.abs(d))


def test_upload_download_round_trip():
    shape = (12, 10, 8)
    n = int(np.prod(shape))
    st = random_state(n, 15, count=9)
    g = gpu_grid(shape, (1, 1, 1), 1e-12, st)
    for c in range(9):
        assert np.array_equal(g.download(c), st[c])
    g32 = gpu_grid(shape, (1, 1, 1), 1e-12, st[:6], Precision.FP32)
    for c in range(6):
        assert np.array_equal(g32.download(c), st[c].astype(np.float32).astype(np.float64))


def test_errors_match_reference():
    g = gpu_grid((4, 4, 4), (1, 1, 1), 1e-12)
    with pytest.raises(LookupError, match="ERROR: Invalid field component"):
        g.get_field(9)
    with pytest.raises(LookupError, match="ERROR: Invalid field component"):
        g.download(11)
    with pytest.raises(_lib.FDTDError):
        g.set_point_sources([dict(i=4, j=0, k=0, comp=2)])
    with pytest.raises(ValueError):
        g.upload(0, np.zeros(10))
    with pytest.raises(_lib.FDTDError):
        g.set_tuning(3, 1, 1)


def test_autotune_keeps_state():
    shape = (64, 32, 32)
    n = int(np.prod(shape))
    st = random_state(n, 16)
    g = gpu_grid(shape, (1, 1, 1), courant_dt(1, 1, 1), st)
    before = g.checksum()
    tune, ms = g.autotune(reps=2)
    assert ms > 0 and tune[0] in (1, 2, 4)
    assert g.checksum() == before
    assert g.step == 0


def test_device_plane_wave_matches_numpy():
    """fdtd_add_plane_wave against the numpy form (Yee positions, B = khat x E)."""
    from fdtd_method_amd.batch import yee_plane_wave
    shape = (24, 20, 16)
    g = gpu_grid(shape, (1, 1, 1), 1e-12)
    pol = [0.3, -1.0, 0.5]
    g.add_plane_wave([2, -1, 3], 0.7, 0.4, pol)
    g.add_plane_wave([0, 0, 1], 0.2, 1.1, [1.0, 0.0, 0.0])
    want = [a + b for a, b in zip(yee_plane_wave(shape, [2, -1, 3], 0.7, 0.4, pol),
                                  yee_plane_wave(shape, [0, 0, 1], 0.2, 1.1, [1.0, 0.0, 0.0]))]
    out = g.fields()
    for c in range(6):
        np.testing.assert_allclose(out[c], want[c], rtol=0, atol=1e-13)
    # E is transverse: the discrete wave stays a wave (energy nearly constant)
    e0 = sum(g.sum_squares())
    g.update_fields(50)
    assert abs(sum(g.sum_squares()) / e0 - 1.0) < 1e-2


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
@pytest.mark.parametrize("L,V,lanes,TYg,kchunk", [(1, 2, 32, 8, 3), (1, 1, 16, 6, 0),
                                                  (2, 2, 64, 0, 0), (2, 1, 32, 12, 5), (3, 1, 64, 8, 0),
                                                  (3, 2, 16, 10, 7), (4, 2, 32, 12, 4), (4, 1, 20, 12, 3)])
def test_time_blocking_bit_exact(prec, L, V, lanes, TYg, kchunk):
    """L steps per pass (ghost-zone tiles) give exactly the unblocked fields,
    with waveform sources, remainders and wrap-around at the tile edges."""
    shape = (40, 36, 22)
    n = int(np.prod(shape))
    dt = courant_dt(1, 1, 1, factor=0.5)
    st = random_state(n, 30)
    src = [dict(i=3, j=35, k=0, comp=2, kind=1, amp=1.0, period=16 * dt, t0=5 * dt, sigma=6 * dt),
           dict(i=39, j=0, k=21, comp=0, kind=0, amp=0.3)]
    a = gpu_grid(shape, (1, 1, 1), dt, st, prec)
    b = gpu_grid(shape, (1, 1, 1), dt, st, prec)
    for g in (a, b):
        g.set_point_sources(src)
    b.set_time_blocking(L, V, lanes, TYg, kchunk)
    print("blocking", b.get_time_blocking(), flush=True)
    for nsteps in (2 * L + 1, L, 3):
        a.update_fields(nsteps)
        b.update_fields(nsteps)
        assert a.checksum() == b.checksum(), nsteps
    if prec == Precision.FP64:
        r = oracle_grid(shape, (1, 1, 1), dt, st)
        r.run_point_sources(src, 2 * L + 1 + L + 3)
        rel = rel_l2(b.fields(), r.E + r.B)
        assert rel <= 1e-12


def test_time_blocking_tiny_grids():
    """Grids smaller than one tile: ghost cells wrap several times."""
    for shape in [(4, 3, 5), (2, 2, 2), (6, 1, 3)]:
        n = int(np.prod(shape))
        dt = courant_dt(1, 1, 1)
        st = random_state(n, 31)
        a = gpu_grid(shape, (1, 1, 1), dt, st)
        b = gpu_grid(shape, (1, 1, 1), dt, st)
        b.set_time_blocking(2, 1, 8, 8, 1)
        a.update_fields(6)
        b.update_fields(6)
        assert a.checksum() == b.checksum(), shape


def test_autotune_blocking_keeps_state():
    shape = (96, 40, 24)
    n = int(np.prod(shape))
    st = random_state(n, 33)
    g = gpu_grid(shape, (1, 1, 1), courant_dt(1, 1, 1), st, Precision.FP32)
    ref = gpu_grid(shape, (1, 1, 1), courant_dt(1, 1, 1), st, Precision.FP32)
    before = g.checksum()
    cfg, ms = g.autotune_blocking(reps=2)
    assert ms > 0 and 0 <= cfg["L"] <= 4
    assert g.checksum() == before
    g.update_fields(9)
    ref.update_fields(9)
    assert g.checksum() == ref.checksum()

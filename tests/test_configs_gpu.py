"""The benchmark workloads at their full sizes and step counts, with the
accuracy bounds each precision path is held to against the FP64 path.

* 512^3, centre Gaussian-modulated Jz source, 1000 steps: FP32 (autotuned
  strip passes) against FP64 (autotuned strip passes).  The FP64 path is
  itself bit-identical to the reference's FP64 fields (the parity tests),
  so this is the FP32 error against the reference method: relative L2 of
  all six components <= 1e-5 (measured 5.2e-6), and each component's
  max-abs error <= 4e-5 of the largest field value (measured at most 1.2e-5,
  Ez next to the source cell).
* 768^3, 8 seeded plane waves plus the centre source, 2000 steps: FP32 and
  FP16 storage (power-of-two scales rescaled in the run) against FP64,
  relative L2 and energy drift.
* 1024x1024x256, 4 seeded Jz sources of distinct periods, 4000 steps: FP32
  temporally blocked passes against FP64.

Each bound is about three times the error measured on an MI355X (printed
by the test), so that an accuracy regression of a few times fails.
"""
import numpy as np
import pytest

from fdtd_method_amd import FDTD, Parameters, Precision
from fdtd_method_amd.workloads import (DT, add_waves, anisotropic_sources, centre_source, energy,
                                       error_summary, precision_waves)

pytestmark = pytest.mark.gpu


def make(shape, prec, sources, waves=(), fp16_interval=100):
    g = FDTD(Parameters(*shape), DT, prec)
    if prec == Precision.FP16:
        g.set_scales([2.0 ** 10] * 6)
    add_waves(g, waves)
    if prec == Precision.FP16:
        if waves:
            g.autoscale(target_log2=12)
            g.set_autoscale_interval(fp16_interval, 12)
        else:
            g.autoscale(target_log2=10, magnitude_floor=4.0 * np.pi * DT * 1e-2)
            g.set_autoscale_interval(20, 10)
    g.set_point_sources(sources)
    cfg, ms = g.autotune_blocking(reps=3)
    print(f"{prec.name} {shape}: blocking {cfg} ({ms:.3f} ms/step)", flush=True)
    return g


def report(name, err):
    print(f"{name}: rel L2 {err['rel_l2']:.3e}, max abs {err['max_abs']:.3e}", flush=True)
    for c, v in err["per_component"].items():
        print(f"   {c}: rel L2 of total {v['rel_l2_of_total']:.3e}  max abs {v['max_abs']:.3e}  "
              f"rms ref {v['rms_ref']:.3e}", flush=True)


def test_headline_512_fp32_vs_fp64_1000_steps():
    n, steps = 512, 1000
    src = [centre_source(n)]
    ref = make((n, n, n), Precision.FP64, src)
    ref.update_fields(steps)
    peak = max(ref.max_abs())
    g = make((n, n, n), Precision.FP32, src)
    g.update_fields(steps)
    err = error_summary(g, ref)
    report("512^3 FP32 vs FP64, 1000 steps", err)
    print(f"   max |field| {peak:.3e}", flush=True)
    assert peak > 0
    assert err["rel_l2"] <= 1e-5
    for c, v in err["per_component"].items():
        assert v["max_abs"] <= 4e-5 * peak, c
    g.close()
    ref.close()


def test_precision_sweep_768_2000_steps():
    n, steps = 768, 2000
    waves = precision_waves(2024)
    src = [centre_source(n)]
    ref = make((n, n, n), Precision.FP64, src, waves)
    e0 = energy(ref)
    ref.update_fields(steps)
    e64 = energy(ref)
    print(f"768^3 FP64 energy {e0:.6e} -> {e64:.6e}", flush=True)
    bounds = {Precision.FP32: (6e-6, 1e-9), Precision.FP16: (3e-2, 3e-4)}
    for prec, (l2_bound, drift_bound) in bounds.items():
        g = make((n, n, n), prec, src, waves)
        g.update_fields(steps)
        err = error_summary(g, ref)
        drift = energy(g) / e64 - 1.0
        report(f"768^3 {prec.name} vs FP64, 2000 steps", err)
        print(f"   energy drift vs FP64 {drift:.3e}", flush=True)
        assert err["rel_l2"] <= l2_bound, prec
        assert abs(drift) <= drift_bound, prec
        g.close()
    ref.close()


def test_anisotropic_1024x1024x256_4000_steps():
    dims, steps = (1024, 1024, 256), 4000
    srcs, periods = anisotropic_sources(dims, 2024)
    print("source periods (dt)", periods, flush=True)
    ref = make(dims, Precision.FP64, srcs)
    ref.update_fields(steps)
    g = make(dims, Precision.FP32, srcs)
    assert g.get_time_blocking_strips()["L"] >= 2  # the temporally blocked passes
    g.update_fields(steps)
    err = error_summary(g, ref)
    report("1024x1024x256 FP32 vs FP64, 4000 steps", err)
    assert err["rel_l2"] <= 4.5e-5
    g.close()
    ref.close()

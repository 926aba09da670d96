"""The reference's unit tests and driver, restated against the drop-in API.

* unit-tests/test_FDTD_method.cpp:71-203 (and the Kokkos copy): 12
  convergence-order assertions, err(N)/err(2N) = 4 +- 0.1, on the Python
  mirror and on the C++ class FDTD_mi355x::FDTD (tests/cpp).
* perf-tests/sample/sample.cpp: the C++ sample driver on the drop-in class,
  its Ex dump bit-identical to the reference class's run (golden SHA-256).
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from fdtd_method_amd import FDTD, Parameters, Precision
from oracle import oracle as O
from test_oracle import CONVERGENCE, run_convergence

pytestmark = pytest.mark.gpu


def gpu_stepper(precision):
    def step(shape, d, dt, st, iters):
        g = FDTD(Parameters(*shape, 0.0, 1.0, 0.0, 2.0, 0.0, 3.0, *d), dt, precision)
        for c in range(6):
            g.upload(c, st[c])
        for _ in range(iters):
            g.update_fields()  # one reference update_fields() per call
        out = g.fields()
        g.close()
        return out
    return step


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
@pytest.mark.parametrize("name,comp,e,b,shape", CONVERGENCE, ids=[c[0] for c in CONVERGENCE])
def test_convergence(name, comp, e, b, shape, precision):
    ratio = run_convergence(gpu_stepper(precision), comp, e, b, shape)
    assert abs(ratio - 4.0) <= 0.1, ratio


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cpp_unit_tests(prec):
    exe = os.path.join(ROOT, "tests", "cpp", "test_fdtd_mi355x")
    res = subprocess.run([exe, prec], capture_output=True, text=True, timeout=300)
    print(res.stdout[-3000:])
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    assert "21 tests, 0 failures" in res.stdout


def test_cpp_sample_driver_matches_reference(tmp_path, golden_dir):
    rec = json.load(open(os.path.join(golden_dir, "ref_sample_32_100.json")))
    exe = os.path.join(ROOT, "fdtd_method_amd", "bin", "sample")
    dump = tmp_path / "ex.bin"
    res = subprocess.run([exe, str(rec["n"]), str(rec["iterations"]), "--dump", str(dump)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    assert res.stdout.startswith("Execution time: ")
    ex = np.fromfile(dump, dtype="<f8")
    assert hashlib.sha256(ex.tobytes()).hexdigest() == rec["sha256"][0]
    lines = res.stdout.strip().splitlines()[1:11]
    printed = [[float(x) for x in ln.split()] for ln in lines]
    np.testing.assert_allclose(printed, rec["ex_slice"], atol=5e-6)
    bad = subprocess.run([exe, "32"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "ERROR: Incorrect number of parameters" in bad.stdout


def test_cpp_sample_driver_pml_block_matches_reference(tmp_path, golden_dir):
    """`sample 32 100 --pml`: the reference driver's __PML_TEST__ block
    (perf-tests/sample/sample.cpp:93-123,136-147), FDTD_PML with 0.2."""
    rec = json.load(open(os.path.join(golden_dir, "ref_pml_sample_32_100.json")))
    plain = json.load(open(os.path.join(golden_dir, "ref_sample_32_100.json")))
    exe = os.path.join(ROOT, "fdtd_method_amd", "bin", "sample")
    dump, dump_pml = tmp_path / "ex.bin", tmp_path / "ex_pml.bin"
    res = subprocess.run([exe, str(rec["n"]), str(rec["iterations"]), "--pml",
                          "--dump", str(dump), "--dump-pml", str(dump_pml)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    assert lines[0].startswith("Execution time: ")
    assert lines[1].startswith("Execution time (PML): ")
    ex = np.fromfile(dump, dtype="<f8")
    assert hashlib.sha256(ex.tobytes()).hexdigest() == plain["sha256"][0]
    ex = np.fromfile(dump_pml, dtype="<f8")
    assert hashlib.sha256(ex.tobytes()).hexdigest() == rec["sha256"][0]
    k = lines.index("PML: ")
    printed = [[float(x) for x in ln.split()] for ln in lines[k + 2:k + 12]]
    np.testing.assert_allclose(printed, rec["ex_slice"], atol=5e-6)


def test_cpp_sample_driver_writes_visualization_slices(tmp_path):
    """`sample N iters 1` (the call python_script_legend/visualization.py:48-49
    makes) writes OutFiles_<1..6>/<t>.csv; read back exactly as that script
    reads them (pandas, sep=';', decimal commas), each slice is plane Nk/2 of
    the oracle's field after t+1 steps."""
    import pandas as pd
    exe = os.path.join(ROOT, "fdtd_method_amd", "bin", "sample")
    n, it = 16, 6
    res = subprocess.run([exe, str(n), str(it), "1"], capture_output=True, text=True,
                         timeout=300, cwd=tmp_path)
    assert res.returncode == 0, res.stdout + res.stderr
    for t in range(it):
        ref = O.spherical_wave(n, t + 1)
        for c in range(6):
            data = pd.read_csv(tmp_path / f"OutFiles_{c + 1}" / f"{t}.csv", sep=";")
            data = data.apply(lambda x: x.str.replace(",", ".").astype(float)
                              if x.dtype == "object" else x)
            assert data.shape == (n, n)
            plane = ref[c].reshape(n, n, n)[n // 2]
            np.testing.assert_allclose(data.to_numpy(dtype=float), plane, rtol=1e-8,
                                       atol=1e-300)
    bad = subprocess.run([exe, "16", "6", "2"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "ERROR: Incorrect number of parameters" in bad.stdout


def test_download_plane_matches_full_download():
    from fdtd_method_amd import FDTD, Parameters, Precision
    rng = np.random.default_rng(3)
    for prec in (Precision.FP64, Precision.FP32):
        g = FDTD(Parameters(12, 10, 7), 1e-12, prec)
        for c in range(6):
            g.upload(c, rng.standard_normal(12 * 10 * 7))
        g.set_current_values([5 + 3 * 12 + 4 * 120], [8], [2.5])
        for c in range(9):
            full = g.download(c).reshape(7, 10, 12)
            for k in (0, 4, 6):
                assert np.array_equal(g.download_plane(c, k), full[k])

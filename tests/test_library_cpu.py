"""Library checks that need no GPU: the C-ABI library loads, exports every
symbol include/fdtd_mi355x.h declares, and rejects invalid parameters with
the reference's messages before touching the device."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from fdtd_method_amd import FDTD, Parameters, _lib

HEADER = os.path.join(ROOT, "include", "fdtd_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_bindings_list():
    assert declared_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing


def test_library_is_gfx950_code():
    """The shared object carries a gfx950 code object (hipcc --offload-arch=gfx950)."""
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_invalid_parameters_without_gpu():
    """src/FDTD/FDTD.cpp:5-7: N <= 0 or dt <= 0 -> invalid_argument."""
    L = _lib.load()
    h = ctypes.c_void_p()
    for dims, dt in (((0, 4, 4), 1e-12), ((4, -1, 4), 1e-12), ((4, 4, 4), 0.0), ((4, 4, 4), -1.0)):
        p = _lib.Parameters(*dims, 0, 1, 0, 1, 0, 1, 1, 1, 1)
        rc = L.fdtd_create(ctypes.byref(p), dt, 0, ctypes.byref(h))
        assert rc == _lib.ERR_INVALID_PARAMETERS
        assert L.fdtd_last_error() == b"ERROR: invalid parameters"
    with pytest.raises(ValueError, match="ERROR: invalid parameters"):
        FDTD(Parameters(4, 4, 0), 1e-12)


def test_null_arguments_are_rejected():
    L = _lib.load()
    assert L.fdtd_create(None, 1.0, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert L.fdtd_update_fields(None, 1) == _lib.ERR_INVALID_ARGUMENT
    assert L.fdtd_destroy(None) == _lib.OK
    # device arrays, the stream and the protected box updates
    ptr, eb, sc = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_double()
    assert L.fdtd_device_field(None, 0, ctypes.byref(ptr), ctypes.byref(eb),
                               ctypes.byref(sc)) == _lib.ERR_INVALID_ARGUMENT
    assert L.fdtd_get_stream(None, ctypes.byref(ptr)) == _lib.ERR_INVALID_ARGUMENT
    box = (ctypes.c_int * 3)(0, 0, 0)
    assert L.fdtd_update_b_box(None, box, box) == _lib.ERR_INVALID_ARGUMENT
    assert L.fdtd_update_e_box(None, box, box) == _lib.ERR_INVALID_ARGUMENT
    assert b"null" in L.fdtd_last_error()


def test_cpp_programs_are_built():
    for path in (os.path.join(ROOT, "fdtd_method_amd", "bin", "sample"),
                 os.path.join(ROOT, "tests", "cpp", "test_fdtd_mi355x")):
        assert os.access(path, os.X_OK), f"{path} missing: run __graft_entry__.build()"


def test_pml_invalid_arguments_without_gpu():
    """fdtd_create_pml validates before touching the device: the reference's
    invalid_argument for bad dimensions, and pml_percent outside [0, 0.5]
    (where the reference's layer slabs would overlap)."""
    from fdtd_method_amd import FDTD_PML
    L = _lib.load()
    h = ctypes.c_void_p()
    p = _lib.Parameters(0, 4, 4, 0, 1, 0, 1, 0, 1, 1, 1, 1)
    assert L.fdtd_create_pml(ctypes.byref(p), 1e-12, 0.2, 0, ctypes.byref(h)) == \
        _lib.ERR_INVALID_PARAMETERS
    p = _lib.Parameters(8, 8, 8, 0, 1, 0, 1, 0, 1, 1, 1, 1)
    for pct in (-0.1, 0.51, float("nan")):
        assert L.fdtd_create_pml(ctypes.byref(p), 1e-12, pct, 0, ctypes.byref(h)) == \
            _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(ValueError, match="ERROR: invalid parameters"):
        FDTD_PML(Parameters(4, 4, 0), 1e-12, 0.2)


@pytest.mark.parametrize("n,pct,d", [(512, 0.2, 1.0), (37, 0.3, 1.7), (10, 0.5, 0.4),
                                     (9, 0.05, 1.0), (64, 0.2, 3e10)])
def test_pml_profile_matches_the_oracle(n, pct, d):
    """The library's layer tables start from this host-side profile
    (fdtd_pml_profile, no GPU): bit-identical to the oracle's sigma, and the
    decay/coefficient follow FDTD_PML.cpp:63-65,98-111."""
    import math
    from fdtd_method_amd import pml_profile
    from oracle import oracle as O
    dt = 0.5 * d / (O.C * math.sqrt(3.0))
    sigma, decay, coef = pml_profile(n, pct, d, dt)
    p = O.pml_sizes(n, n, n, pct)[0]
    ref = O.pml_sigma(n, p, d)
    assert (sigma == ref).all()
    for i in range(n):
        a = math.exp(-ref[i] * dt * O.C)
        assert decay[i] == a
        c = (1.0 - math.exp(-ref[i] * dt * O.C)) / (ref[i] * d) if ref[i] != 0.0 else O.C * dt / d
        assert coef[i] == c

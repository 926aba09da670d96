"""Build recipes for the native parts of fdtd_method_amd.

``build_library()`` compiles the HIP sources under ``csrc/`` for gfx950 into
``fdtd_method_amd/libfdtd_mi355x.so`` (in-tree, so it travels with the
source tree).  Every call compiles; nothing is cached.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libfdtd_mi355x.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)

HIP_SOURCES = ["fdtd_runtime.hip", "fdtd_batch.hip"]

# -ffp-contract=off keeps every a*b+c unfused, as in the reference's CPU
# build (no FMA): the FP64 path is then bit-identical to the reference.
# -packed-fp32-ops keeps the compiler from emitting packed FP32 math.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# roctx ranges (rocprofv3 --marker-trace) come from the ROCm roctx library.
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined",
              "-lrocprofiler-sdk-roctx"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False, extra_flags=None) -> str:
    """Compile the HIP library for gfx950 (one hipcc per source, run
    concurrently, then one link); returns the .so path."""
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(PKG, src.replace(".hip", ".o.tmp"))
        cmd = [hipcc(), *HIPCC_FLAGS, *(extra_flags or []), f"-I{INCLUDE}", f"-I{CSRC}", "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    failed = []
    for src, proc in procs:
        _, err = proc.communicate()
        if proc.returncode != 0:
            failed.append(f"hipcc failed for {src} ({proc.returncode}):\n{err[-8000:]}")
    try:
        if failed:
            raise RuntimeError("\n".join(failed))
        tmp = LIB_PATH + ".tmp"
        cmd = [hipcc(), *LINK_FLAGS, *objs, "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed ({res.returncode}):\n{res.stderr[-8000:]}")
        os.replace(tmp, LIB_PATH)
    finally:
        for obj in objs:
            if os.path.exists(obj):
                os.remove(obj)
    return LIB_PATH


BIN_DIR = os.path.join(PKG, "bin")
CPP_TEST_DIR = os.path.join(ROOT, "tests", "cpp")
CXX_FLAGS = ["-O2", "-std=c++17", "-Wall", "-Wno-unused-function"]


def _gxx(src: str, out: str, rpath: str, verbose: bool) -> None:
    cmd = ["g++", *CXX_FLAGS, f"-I{INCLUDE}", src, "-o", out, f"-L{PKG}", "-lfdtd_mi355x",
           f"-Wl,-rpath,{rpath}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{res.stderr[-6000:]}")


def _hipcc_program(src: str, out: str, rpath: str, verbose: bool) -> None:
    cmd = [hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", f"-I{INCLUDE}", src, "-o", out, f"-L{PKG}",
           "-lfdtd_mi355x", f"-Wl,-rpath,{rpath}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")


def build_cpp_tests(verbose: bool = False) -> None:
    """Compile the programs that use the drop-in class: the sample driver
    (fdtd_method_amd/bin/sample), the device-current driver
    (fdtd_method_amd/bin/device_sample, HIP: it writes J in a kernel of its
    own) and the C++ unit test (tests/cpp/test_fdtd_mi355x).  All link
    libfdtd_mi355x.so."""
    os.makedirs(BIN_DIR, exist_ok=True)
    _gxx(os.path.join(CSRC, "sample.cpp"), os.path.join(BIN_DIR, "sample"), "$ORIGIN/..", verbose)
    _hipcc_program(os.path.join(CSRC, "device_sample.hip"), os.path.join(BIN_DIR, "device_sample"),
                   "$ORIGIN/..", verbose)
    _gxx(os.path.join(CPP_TEST_DIR, "test_fdtd_mi355x.cpp"),
         os.path.join(CPP_TEST_DIR, "test_fdtd_mi355x"), "$ORIGIN/../../fdtd_method_amd", verbose)


def build_oracle(verbose: bool = False) -> None:
    """Compile the CPU checker (oracle/liboracle.so) and, when the reference
    checkout exists, oracle/_ref/libfdtd_ref.so and the reference's own
    sample driver (oracle/_ref/sample, sample_pml) for CPU timings."""
    odir = os.path.join(ROOT, "oracle")
    targets = ["oracle"]
    ref = os.environ.get("FDTD_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src", "FDTD")):
        targets += ["ref", "ref-sample"]
    cmd = ["make", "-B", "-C", odir, f"REFERENCE={ref}", *targets]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose:
        print(res.stdout)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout[-4000:]}\n{res.stderr[-4000:]}")

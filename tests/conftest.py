"""Test configuration: the ``gpu`` marker and shared helpers.

``-m "not gpu"`` runs the oracle pins, golden fixtures and library checks on
the CPU; ``-m gpu`` runs the parity tests through the C-ABI on an MI355X.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    config.addinivalue_line("markers", "torch: uses torch's GPU tensors")


def pytest_collection_modifyitems(config, items):
    # torch ships its own HIP runtime; started after this library has run
    # kernels in the process it reports no device.  GPU tests that hand the
    # library's arrays to torch therefore start torch before any grid exists.
    if "not gpu" in (config.option.markexpr or ""):
        return
    if any(item.get_closest_marker("torch") for item in items):
        try:
            import torch
            torch.cuda.is_available()
        except Exception:
            pass


def random_state(n, seed, count=6):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(n) for _ in range(count)]


def courant_dt(dx, dy, dz, factor=0.45):
    from oracle.oracle import C
    return factor / (C * np.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN

"""MI355X (gfx950) FDTD Yee time step behind the reference's FDTD class API.

The compute path is libfdtd_mi355x.so (hand-written HIP kernels, C-ABI in
include/fdtd_mi355x.h); this package holds its Python bindings and the
host-side mirror of the reference interface.
"""
from .structures import (C, PI, Axis, Component, CurrentParameters, Parameters, Precision,
                         SelectedFields)
from .fdtd import FDTD, FDTD_PML, Field, calibration_copy, pml_profile, set_tune_cache
from .batch import Batch, plane_wave_batch

__all__ = ["C", "PI", "Axis", "Component", "CurrentParameters", "Parameters", "Precision",
           "SelectedFields", "FDTD", "FDTD_PML", "Field", "calibration_copy", "pml_profile",
           "set_tune_cache", "Batch", "plane_wave_batch"]

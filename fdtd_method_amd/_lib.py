"""ctypes bindings of libfdtd_mi355x.so (C-ABI: include/fdtd_mi355x.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libfdtd_mi355x.so")

OK = 0
ERR_INVALID_PARAMETERS = -1
ERR_INVALID_COMPONENT = -2
ERR_INVALID_ARGUMENT = -3
ERR_HIP = -4
ERR_UNSUPPORTED = -5

FP64, FP32, FP16 = 0, 1, 2
PATH_FUSED, PATH_3SWEEP = 0, 1

# Every symbol include/fdtd_mi355x.h declares (checked by the CPU test suite).
EXPORTED = [
    "fdtd_last_error", "fdtd_create", "fdtd_destroy", "fdtd_create_pml", "fdtd_get_pml",
    "fdtd_set_pml_path", "fdtd_get_pml_path",
    "fdtd_pml_download_split", "fdtd_pml_profile", "fdtd_upload", "fdtd_download",
    "fdtd_download_plane",
    "fdtd_zero_currents", "fdtd_set_current_values", "fdtd_set_point_sources",
    "fdtd_device_field", "fdtd_get_stream", "fdtd_update_b_box", "fdtd_update_e_box",
    "fdtd_update_fields_with_currents", "fdtd_get_stats",
    "fdtd_set_probes", "fdtd_probe_rows", "fdtd_read_probes", "fdtd_set_energy_interval",
    "fdtd_energy_rows", "fdtd_read_energy", "fdtd_rewind_records",
    "fdtd_update_fields", "fdtd_update_fields_3sweep", "fdtd_get_step", "fdtd_set_step",
    "fdtd_synchronize", "fdtd_get_info", "fdtd_sum_squares", "fdtd_sum_squares_noalloc",
    "fdtd_max_abs", "fdtd_compare", "fdtd_count_nonfinite",
    "fdtd_checksum", "fdtd_add_plane_wave",
    "fdtd_set_scales", "fdtd_get_scales", "fdtd_autoscale", "fdtd_set_autoscale_interval",
    "fdtd_zero_fields", "fdtd_set_step_carry", "fdtd_set_time_blocking", "fdtd_get_time_blocking",
    "fdtd_autotune_blocking", "fdtd_set_tune_cache", "fdtd_set_time_blocking_strips", "fdtd_get_time_blocking_strips",
    "fdtd_set_tuning",
    "fdtd_get_tuning", "fdtd_autotune", "fdtd_time_update_fields", "fdtd_time_update_fields_kernels",
    "fdtd_calibration_copy",
    "fdtd_batch_create", "fdtd_batch_destroy", "fdtd_batch_update_fields",
    "fdtd_batch_update_fields_graph", "fdtd_batch_synchronize", "fdtd_batch_time",
    "fdtd_batch_set_shape", "fdtd_batch_get_shape", "fdtd_batch_autotune",
    "fdtd_batch_set_time_blocking_strips", "fdtd_batch_get_time_blocking_strips",
]


class Parameters(ctypes.Structure):
    """include/Structures.h:23-38 (FDTD_struct::Parameters)."""
    _fields_ = [("Ni", ctypes.c_int), ("Nj", ctypes.c_int), ("Nk", ctypes.c_int),
                ("ax", ctypes.c_double), ("bx", ctypes.c_double),
                ("ay", ctypes.c_double), ("by", ctypes.c_double),
                ("az", ctypes.c_double), ("bz", ctypes.c_double),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class PointSource(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int), ("j", ctypes.c_int), ("k", ctypes.c_int),
                ("comp", ctypes.c_int), ("kind", ctypes.c_int), ("reserved", ctypes.c_int),
                ("amp", ctypes.c_double), ("period", ctypes.c_double), ("t0", ctypes.c_double),
                ("sigma", ctypes.c_double), ("phase", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_lib = None


class FDTDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def load():
    """Load the HIP library (built by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "fdtd_last_error": ([], ctypes.c_char_p),
        "fdtd_create": ([ctypes.POINTER(Parameters), ctypes.c_double, _i, ctypes.POINTER(_vp)], _i),
        "fdtd_destroy": ([_vp], _i),
        "fdtd_create_pml": ([ctypes.POINTER(Parameters), ctypes.c_double, ctypes.c_double, _i,
                             ctypes.POINTER(_vp)], _i),
        "fdtd_get_pml": ([_vp, ctypes.POINTER(_i), _dp], _i),
        "fdtd_set_pml_path": ([_vp, _i, _i, _i, _i], _i),
        "fdtd_get_pml_path": ([_vp, ctypes.POINTER(_i)], _i),
        "fdtd_pml_download_split": ([_vp, _i, _dp, ctypes.c_size_t], _i),
        "fdtd_pml_profile": ([_i, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp,
                              _dp], _i),
        "fdtd_upload": ([_vp, _i, _dp, ctypes.c_size_t], _i),
        "fdtd_download": ([_vp, _i, _dp, ctypes.c_size_t], _i),
        "fdtd_download_plane": ([_vp, _i, _i, _dp, ctypes.c_size_t], _i),
        "fdtd_zero_currents": ([_vp], _i),
        "fdtd_device_field": ([_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_i),
                               ctypes.POINTER(ctypes.c_double)], _i),
        "fdtd_get_stream": ([_vp, ctypes.POINTER(_vp)], _i),
        "fdtd_update_b_box": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)], _i),
        "fdtd_update_e_box": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)], _i),
        "fdtd_set_current_values": ([_vp, _i, ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.POINTER(_i), _dp], _i),
        "fdtd_set_point_sources": ([_vp, _i, ctypes.POINTER(PointSource)], _i),
        "fdtd_update_fields_with_currents": ([_vp, _i, ctypes.POINTER(_i),
                                              ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_i),
                                              _dp], _i),
        "fdtd_get_stats": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_set_probes": ([_vp, _i, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_i),
                             ctypes.c_longlong], _i),
        "fdtd_probe_rows": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_read_probes": ([_vp, ctypes.c_longlong, ctypes.c_longlong, _dp], _i),
        "fdtd_set_energy_interval": ([_vp, _i, ctypes.c_longlong], _i),
        "fdtd_energy_rows": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_read_energy": ([_vp, ctypes.c_longlong, ctypes.c_longlong,
                              ctypes.POINTER(ctypes.c_longlong), _dp], _i),
        "fdtd_rewind_records": ([_vp], _i),
        "fdtd_update_fields": ([_vp, _i], _i),
        "fdtd_update_fields_3sweep": ([_vp, _i], _i),
        "fdtd_get_step": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_set_step": ([_vp, ctypes.c_longlong], _i),
        "fdtd_synchronize": ([_vp], _i),
        "fdtd_get_info": ([_vp, ctypes.POINTER(_i), ctypes.POINTER(_i),
                           ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_sum_squares": ([_vp, _dp], _i),
        "fdtd_sum_squares_noalloc": ([_vp, _dp], _i),
        "fdtd_max_abs": ([_vp, _dp], _i),
        "fdtd_count_nonfinite": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_compare": ([_vp, _vp, _dp], _i),
        "fdtd_checksum": ([_vp, ctypes.POINTER(ctypes.c_ulonglong)], _i),
        "fdtd_add_plane_wave": ([_vp, ctypes.POINTER(_i), ctypes.c_double, ctypes.c_double,
                                 _dp], _i),
        "fdtd_set_scales": ([_vp, _dp], _i),
        "fdtd_get_scales": ([_vp, _dp], _i),
        "fdtd_autoscale": ([_vp, _i, ctypes.c_double], _i),
        "fdtd_set_autoscale_interval": ([_vp, _i, _i], _i),
        "fdtd_zero_fields": ([_vp], _i),
        "fdtd_set_step_carry": ([_vp, _i], _i),
        "fdtd_set_time_blocking": ([_vp, _i, _i, _i, _i, _i], _i),
        "fdtd_get_time_blocking": ([_vp, ctypes.POINTER(_i)], _i),
        "fdtd_autotune_blocking": ([_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_set_tune_cache": ([ctypes.c_char_p], _i),
        "fdtd_set_time_blocking_strips": ([_vp, _i, _i, _i, _i, _i], _i),
        "fdtd_get_time_blocking_strips": ([_vp, ctypes.POINTER(_i)], _i),
        "fdtd_set_tuning": ([_vp, _i, _i, _i], _i),
        "fdtd_get_tuning": ([_vp, ctypes.POINTER(_i)], _i),
        "fdtd_autotune": ([_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_time_update_fields": ([_vp, _i, _i, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_time_update_fields_kernels": ([_vp, _i, _dp, ctypes.POINTER(_i)], _i),
        "fdtd_calibration_copy": ([ctypes.c_size_t, _i, ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_batch_create": ([_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], _i),
        "fdtd_batch_destroy": ([_vp], _i),
        "fdtd_batch_update_fields": ([_vp, _i], _i),
        "fdtd_batch_update_fields_graph": ([_vp, _i, _i], _i),
        "fdtd_batch_synchronize": ([_vp], _i),
        "fdtd_batch_time": ([_vp, _i, ctypes.POINTER(ctypes.c_float),
                             ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_batch_set_shape": ([_vp, _i, _i, ctypes.c_longlong], _i),
        "fdtd_batch_get_shape": ([_vp, ctypes.POINTER(ctypes.c_longlong)], _i),
        "fdtd_batch_autotune": ([_vp, _i, ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_float)], _i),
        "fdtd_batch_set_time_blocking_strips": ([_vp, _i, _i, _i, _i], _i),
        "fdtd_batch_get_time_blocking_strips": ([_vp, ctypes.POINTER(_i)], _i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        msg = load().fdtd_last_error().decode(errors="replace")
        if rc == ERR_INVALID_PARAMETERS:
            raise ValueError(msg)
        if rc == ERR_INVALID_COMPONENT:
            raise LookupError(msg)
        raise FDTDError(rc, msg)

"""ctypes binding of libdsmvc_hip.so (the C ABI declared in include/dsmvc_hip.h).

Loading never falls back to anything: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdsmvc_hip.so"

OK, INVALID_ARGUMENT, RUNTIME_ERROR, UNAVAILABLE = 0, 1, 2, 3
ASYNC = 1

# Every symbol include/dsmvc_hip.h declares.
EXPORTS = (
    "dsmvc_hip_last_error", "dsmvc_hip_plan_build", "dsmvc_hip_plan_build_many",
    "dsmvc_hip_plan_release",
    "dsmvc_hip_plan_info_get", "dsmvc_hip_plan_arrays", "dsmvc_hip_plan_window_info",
    "dsmvc_hip_plan_operator",
    "dsmvc_hip_planner_cache_stats",
    "dsmvc_hip_clear_planner_caches", "dsmvc_hip_padding_index", "dsmvc_hip_parse_backend",
    "dsmvc_hip_resolve_backend", "dsmvc_hip_device_count", "dsmvc_hip_executor_create",
    "dsmvc_hip_executor_destroy", "dsmvc_hip_executor_prepare", "dsmvc_hip_executor_seal",
    "dsmvc_hip_executor_tune", "dsmvc_hip_executor_set_refinement",
    "dsmvc_hip_executor_tuned_route", "dsmvc_hip_inverse_rows",
    "dsmvc_hip_inverse_columns",
    "dsmvc_hip_inverse_2d", "dsmvc_hip_inverse_2d_u8", "dsmvc_hip_inverse_2d_u16",
    "dsmvc_hip_inverse_2d_batch", "dsmvc_hip_inverse_rows_batch",
    "dsmvc_hip_inverse_columns_batch", "dsmvc_hip_inverse_2d_planes",
    "dsmvc_hip_batch_create", "dsmvc_hip_batch_run",
    "dsmvc_hip_batch_info_get", "dsmvc_hip_batch_destroy", "dsmvc_hip_server_create",
    "dsmvc_hip_server_submit", "dsmvc_hip_server_wait", "dsmvc_hip_server_poll",
    "dsmvc_hip_server_release", "dsmvc_hip_server_stats_get", "dsmvc_hip_server_destroy",
    "dsmvc_hip_executor_release_stream", "dsmvc_hip_synchronize",
    "dsmvc_hip_inverse_2d_batch_noalloc", "dsmvc_hip_plan_import",
    "dsmvc_hip_inverse_rows_host", "dsmvc_hip_inverse_columns_host", "dsmvc_hip_inverse_2d_host",
    "dsmvc_hip_executor_set_plan_cache", "dsmvc_hip_executor_plan_cache_stats",
    "dsmvc_hip_plan_build_custom", "dsmvc_hip_inverse_2d_exported",
    "dsmvc_hip_inverse_rows_exported", "dsmvc_hip_inverse_columns_exported",
    "dsmvc_hip_import_exported_semaphore", "dsmvc_hip_has_exported_semaphore",
    "dsmvc_hip_close_exported_plane", "dsmvc_hip_executor_import_stats",
)
KERNEL_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)
BATCH_GRAPH = 2


class AxisRequestC(ctypes.Structure):
    _fields_ = [
        ("source_size", ctypes.c_int32),
        ("destination_size", ctypes.c_int32),
        ("active_length", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("kernel", ctypes.c_int32),
        ("taps", ctypes.c_int32),
        ("b", ctypes.c_double),
        ("c", ctypes.c_double),
        ("blur", ctypes.c_double),
        ("border", ctypes.c_int32),
        ("f64_mode", ctypes.c_int32),
    ]


class ConversionC(ctypes.Structure):
    _fields_ = [
        ("input_offset", ctypes.c_float),
        ("input_scale", ctypes.c_float),
        ("output_scale", ctypes.c_float),
        ("output_offset", ctypes.c_float),
        ("output_maximum", ctypes.c_uint32),
    ]


class WindowInfoC(ctypes.Structure):
    _fields_ = [
        ("feasible", ctypes.c_int32),
        ("rows_feasible", ctypes.c_int32),
        ("band", ctypes.c_int32),
        ("chunks", ctypes.c_int32),
        ("length", ctypes.c_int32),
        ("warmup", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("float64_feasible", ctypes.c_int32),
        ("residual", ctypes.c_double),
    ]


class OperatorInfoC(ctypes.Structure):
    _fields_ = [
        ("rows_feasible", ctypes.c_int32),
        ("columns_feasible", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("max_band", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("values", ctypes.c_int64),
        ("dropped", ctypes.c_double),
        ("max_abs", ctypes.c_double),
    ]


class PlaneJobC(ctypes.Structure):
    _fields_ = [
        ("horizontal", ctypes.c_void_p),
        ("vertical", ctypes.c_void_p),
        ("input", ctypes.c_void_p),
        ("input_row_stride", ctypes.c_ssize_t),
        ("input_frame_stride", ctypes.c_ssize_t),
        ("output", ctypes.c_void_p),
        ("output_row_stride", ctypes.c_ssize_t),
        ("output_frame_stride", ctypes.c_ssize_t),
        ("frames", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("conversion", ctypes.c_void_p),
    ]


class BatchJobC(ctypes.Structure):
    _fields_ = [
        ("horizontal", ctypes.c_void_p),
        ("vertical", ctypes.c_void_p),
        ("input", ctypes.c_void_p),
        ("input_row_stride", ctypes.c_ssize_t),
        ("input_frame_stride", ctypes.c_ssize_t),
        ("output", ctypes.c_void_p),
        ("output_row_stride", ctypes.c_ssize_t),
        ("output_frame_stride", ctypes.c_ssize_t),
        ("frames", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class BatchInfoC(ctypes.Structure):
    _fields_ = [
        ("jobs", ctypes.c_int32),
        ("table_jobs", ctypes.c_int32),
        ("fallback_jobs", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("items", ctypes.c_int64),
        ("intermediate_bytes", ctypes.c_int64),
        ("graph", ctypes.c_int32),
        ("fused_jobs", ctypes.c_int32),
    ]


class ServeRequestC(ctypes.Structure):
    _fields_ = [
        ("planes", ctypes.c_int32),
        ("subsampling_w", ctypes.c_int32),
        ("subsampling_h", ctypes.c_int32),
        ("src_width", ctypes.c_int32),
        ("src_height", ctypes.c_int32),
        ("dst_width", ctypes.c_int32),
        ("dst_height", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("taps", ctypes.c_int32),
        ("border", ctypes.c_int32),
        ("f64_mode", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("b", ctypes.c_double),
        ("c", ctypes.c_double),
        ("blur", ctypes.c_double),
        ("src", ctypes.c_void_p * 3),
        ("src_stride", ctypes.c_ssize_t * 3),
        ("dst", ctypes.c_void_p * 3),
        ("dst_stride", ctypes.c_ssize_t * 3),
    ]


class ServerConfigC(ctypes.Structure):
    _fields_ = [
        ("memory_budget", ctypes.c_int64),
        ("slots", ctypes.c_int32),
        ("max_round_requests", ctypes.c_int32),
        ("fused", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServerStatsC(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("rounds", ctypes.c_uint64),
        ("table_planes", ctypes.c_uint64),
        ("fallback_planes", ctypes.c_uint64),
        ("largest_round", ctypes.c_uint64),
        ("memory_high_water", ctypes.c_int64),
        ("fused_planes", ctypes.c_uint64),
    ]


class PlanInfoC(ctypes.Structure):
    _fields_ = [
        ("source_size", ctypes.c_int32),
        ("destination_size", ctypes.c_int32),
        ("support", ctypes.c_int32),
        ("half_bandwidth", ctypes.c_int32),
        ("active_length", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("normal_rcond", ctypes.c_double),
        ("normal_inf_norm", ctypes.c_double),
        ("requires_float64", ctypes.c_int32),
        ("nnz", ctypes.c_int32),
    ]


PLAN_ARRAYS = ("offsets", "indices", "weights", "lower_ld", "upper_l", "inverse_diagonal",
               "weights_f64", "normal_bands_f64", "ldlt_bands_f64", "inverse_diagonal_f64")


class PlanDataC(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in PLAN_ARRAYS]


class PlanCacheStatsC(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64)
                for name in ("entries", "bytes", "limit", "hits", "misses", "evictions")]


class PlannerStatsC(ctypes.Structure):
    _fields_ = [
        ("plan_hits", ctypes.c_uint64),
        ("plan_builds", ctypes.c_uint64),
        ("plan_entries", ctypes.c_uint64),
        ("plan_resident_bytes", ctypes.c_uint64),
        ("geometry_hits", ctypes.c_uint64),
        ("geometry_builds", ctypes.c_uint64),
        ("geometry_entries", ctypes.c_uint64),
        ("geometry_resident_bytes", ctypes.c_uint64),
    ]


class DsmvcError(RuntimeError):
    pass


_lib = None


class ExportedPlaneC(ctypes.Structure):
    _fields_ = [
        ("memory_id", ctypes.c_uint64),
        ("memory_size", ctypes.c_uint64),
        ("offset", ctypes.c_uint64),
        ("plane_bytes", ctypes.c_uint64),
        ("handle_type", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("handle", ctypes.c_int64),
        ("pitch_bytes", ctypes.c_int64),
        ("ready_semaphore", ctypes.c_uint64),
        ("ready_value", ctypes.c_uint64),
    ]


class ImportStatsC(ctypes.Structure):
    _fields_ = [
        ("entries", ctypes.c_uint64),
        ("imports", ctypes.c_uint64),
        ("hits", ctypes.c_uint64),
        ("evictions", ctypes.c_uint64),
    ]


def load():
    """Loads libdsmvc_hip.so (built by `make -C vapoursynth-descalemvc_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          f"make -C vapoursynth-descalemvc_amd")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, pd, i32, u32 = ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int32, ctypes.c_uint32
    lib.dsmvc_hip_last_error.restype = ctypes.c_char_p
    lib.dsmvc_hip_plan_build.argtypes = [ctypes.POINTER(AxisRequestC), ctypes.POINTER(vp)]
    lib.dsmvc_hip_plan_build_many.argtypes = [ctypes.POINTER(AxisRequestC), i32, i32,
                                              ctypes.POINTER(vp)]
    lib.dsmvc_hip_plan_release.argtypes = [vp]
    lib.dsmvc_hip_plan_release.restype = None
    lib.dsmvc_hip_plan_info_get.argtypes = [vp, ctypes.POINTER(PlanInfoC)]
    lib.dsmvc_hip_plan_arrays.argtypes = [vp] + [vp] * 10
    lib.dsmvc_hip_plan_window_info.argtypes = [vp, ctypes.POINTER(WindowInfoC)]
    lib.dsmvc_hip_plan_operator.argtypes = [vp, ctypes.POINTER(OperatorInfoC), vp, vp, vp,
                                            ctypes.c_int64]
    lib.dsmvc_hip_planner_cache_stats.argtypes = [ctypes.POINTER(PlannerStatsC)]
    lib.dsmvc_hip_clear_planner_caches.restype = None
    lib.dsmvc_hip_padding_index.argtypes = [ctypes.c_int64, i32, i32, ctypes.POINTER(i32)]
    lib.dsmvc_hip_parse_backend.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
    lib.dsmvc_hip_resolve_backend.argtypes = [i32, ctypes.POINTER(i32)]
    lib.dsmvc_hip_device_count.argtypes = [ctypes.POINTER(i32)]
    lib.dsmvc_hip_executor_create.argtypes = [i32, i32, ctypes.POINTER(vp)]
    lib.dsmvc_hip_executor_destroy.argtypes = [vp]
    lib.dsmvc_hip_executor_destroy.restype = None
    lib.dsmvc_hip_executor_prepare.argtypes = [vp, vp]
    lib.dsmvc_hip_executor_seal.argtypes = [vp]
    lib.dsmvc_hip_executor_tune.argtypes = [vp, i32, i32]
    lib.dsmvc_hip_executor_set_refinement.argtypes = [vp, i32]
    lib.dsmvc_hip_executor_tuned_route.argtypes = [vp, vp, i32, i32, ctypes.POINTER(i32)]
    for name in ("dsmvc_hip_inverse_rows", "dsmvc_hip_inverse_columns"):
        getattr(lib, name).argtypes = [vp, vp, vp, pd, vp, pd, i32, vp, u32]
    lib.dsmvc_hip_inverse_2d.argtypes = [vp, vp, vp, vp, pd, vp, pd, vp, u32]
    for name in ("dsmvc_hip_inverse_2d_u8", "dsmvc_hip_inverse_2d_u16"):
        getattr(lib, name).argtypes = [vp, vp, vp, vp, pd, vp, pd, ctypes.POINTER(ConversionC),
                                       vp, u32]
    lib.dsmvc_hip_inverse_2d_batch.argtypes = [vp, vp, vp, i32, vp, pd, pd, vp, pd, pd, i32,
                                               ctypes.POINTER(ConversionC), vp, u32]
    lib.dsmvc_hip_plan_build_custom.argtypes = [ctypes.POINTER(AxisRequestC), KERNEL_FN, vp,
                                                ctypes.POINTER(vp)]
    lib.dsmvc_hip_inverse_rows_host.argtypes = [vp, vp, vp, pd, vp, pd, i32]
    lib.dsmvc_hip_inverse_columns_host.argtypes = [vp, vp, vp, pd, vp, pd, i32]
    lib.dsmvc_hip_inverse_2d_host.argtypes = [vp, vp, vp, i32, vp, pd, pd, vp, pd, pd, i32,
                                              ctypes.POINTER(ConversionC)]
    lib.dsmvc_hip_executor_set_plan_cache.argtypes = [vp, ctypes.c_int64]
    lib.dsmvc_hip_executor_plan_cache_stats.argtypes = [vp, ctypes.POINTER(PlanCacheStatsC)]
    lib.dsmvc_hip_plan_import.argtypes = [ctypes.POINTER(PlanInfoC), ctypes.POINTER(PlanDataC),
                                          ctypes.POINTER(vp)]
    lib.dsmvc_hip_inverse_2d_batch_noalloc.argtypes = [vp, vp, vp, i32, vp, pd, pd, vp, pd, pd,
                                                       i32, ctypes.POINTER(ConversionC), vp, u32]
    lib.dsmvc_hip_inverse_rows_batch.argtypes = [vp, vp, i32, vp, pd, pd, vp, pd, pd, i32, i32,
                                                 ctypes.POINTER(ConversionC), vp, u32]
    lib.dsmvc_hip_inverse_columns_batch.argtypes = [vp, vp, vp, pd, pd, i32, vp, pd, pd, i32,
                                                    i32, ctypes.POINTER(ConversionC), vp, u32]
    lib.dsmvc_hip_inverse_2d_planes.argtypes = [vp, ctypes.POINTER(PlaneJobC), i32, i32, vp, u32]
    lib.dsmvc_hip_batch_create.argtypes = [vp, ctypes.POINTER(BatchJobC), i32, u32,
                                           ctypes.POINTER(vp)]
    lib.dsmvc_hip_batch_run.argtypes = [vp, vp, u32]
    lib.dsmvc_hip_batch_info_get.argtypes = [vp, ctypes.POINTER(BatchInfoC)]
    lib.dsmvc_hip_batch_destroy.argtypes = [vp]
    lib.dsmvc_hip_batch_destroy.restype = None
    lib.dsmvc_hip_server_create.argtypes = [vp, ctypes.POINTER(ServerConfigC), ctypes.POINTER(vp)]
    lib.dsmvc_hip_server_submit.argtypes = [vp, ctypes.POINTER(ServeRequestC),
                                            ctypes.POINTER(ctypes.c_uint64)]
    lib.dsmvc_hip_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
    lib.dsmvc_hip_server_poll.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(i32)]
    lib.dsmvc_hip_server_release.argtypes = [vp, ctypes.c_uint64]
    lib.dsmvc_hip_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStatsC)]
    lib.dsmvc_hip_server_destroy.argtypes = [vp]
    lib.dsmvc_hip_server_destroy.restype = None
    xp = ctypes.POINTER(ExportedPlaneC)
    lib.dsmvc_hip_inverse_2d_exported.argtypes = [vp, vp, vp, i32, xp, xp,
                                                  ctypes.POINTER(ConversionC)]
    lib.dsmvc_hip_inverse_rows_exported.argtypes = [vp, vp, xp, xp, i32]
    lib.dsmvc_hip_inverse_columns_exported.argtypes = [vp, vp, xp, xp, i32]
    lib.dsmvc_hip_import_exported_semaphore.argtypes = [vp, ctypes.c_uint64, i32, ctypes.c_int64,
                                                        ctypes.POINTER(i32)]
    lib.dsmvc_hip_has_exported_semaphore.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(i32)]
    lib.dsmvc_hip_close_exported_plane.argtypes = [xp]
    lib.dsmvc_hip_close_exported_plane.restype = None
    lib.dsmvc_hip_executor_import_stats.argtypes = [vp, ctypes.POINTER(ImportStatsC)]
    lib.dsmvc_hip_executor_release_stream.argtypes = [vp, vp]
    lib.dsmvc_hip_synchronize.argtypes = [vp, vp]
    _lib = lib
    return lib


def check(status: int):
    """Maps a C-ABI status to the reference's exception kinds."""
    if status == OK:
        return
    message = load().dsmvc_hip_last_error().decode()
    if status == INVALID_ARGUMENT:
        raise ValueError(message)
    raise DsmvcError(message)

"""dsmvc_hip -- host side of the MI355X descale executor.

Mirrors the reference plugin's operator surface (core.dsmvc.Descale,
Debilinear, Debicubic, Delanczos, Despline16/36/64 with backend="hip";
reference src/vs_plugin.cpp:1989-2068, dsmvc.py:116-150) for planes held as
device tensors, and exposes the engine level (plans, executor) of
include/dsmvc_hip.h.  Every compute call goes through libdsmvc_hip.so; there
is no CPU path here -- without the library or a HIP device the calls raise.

Tensors are torch tensors on a HIP device (torch is only the allocator and
stream provider here).  A clip is a `Clip`: a list of planes, each
[frames, height, width] (or [height, width]) of float32, uint8 or uint16.
"""
from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ASYNC, AxisRequestC, ConversionC, DsmvcError, PlanInfoC, check, load

__all__ = [
    "AxisRequest", "Plan", "Executor", "Clip", "IntegerConversion", "build_plan", "build_plans",
    "filter_requests", "import_plan", "build_custom_plan", "CUSTOM",
    "integer_conversion", "Descale", "Debilinear", "Debicubic", "Delanczos", "Despline16",
    "Despline36", "Despline64", "parse_backend", "resolve_backend", "device_count",
    "padding_index", "planner_cache_stats", "clear_planner_caches", "DsmvcError",
    "BILINEAR", "BICUBIC", "LANCZOS", "SPLINE16", "SPLINE36", "SPLINE64",
    "ZERO", "REPEAT", "REFLECT101", "SYMMETRIC", "MIRROR",
]

# Reference numbering (include/dsmvc/engine.hpp:16-58).
BILINEAR, BICUBIC, LANCZOS, SPLINE16, SPLINE36, SPLINE64 = range(6)
ZERO, REPEAT, REFLECT101, SYMMETRIC, MIRROR = range(5)
KERNELS = {"bilinear": BILINEAR, "bicubic": BICUBIC, "lanczos": LANCZOS,
           "spline16": SPLINE16, "spline36": SPLINE36, "spline64": SPLINE64}
BACKENDS = {0: "auto", 1: "cpu", 2: "metal", 3: "vulkan", 4: "cuda", 5: "hip"}
SOLVE_MODES = {"auto": 0, "exact": 1, "chunked": 2, "windowed": 3, "refined": 4, "fused": 5}
SAMPLE_KINDS = {"float32": 0, "uint8": 1, "uint16": 2}


@dataclass(frozen=True)
class AxisRequest:
    """dsmvc::AxisRequest (reference include/dsmvc/engine.hpp:60-78)."""
    source_size: int
    destination_size: int
    active_length: float
    shift: float = 0.0
    kernel: int = BICUBIC
    taps: int = 3
    b: float = 0.0
    c: float = 0.5
    blur: float = 1.0
    border: int = SYMMETRIC
    f64_mode: int = 0

    def to_c(self) -> AxisRequestC:
        return AxisRequestC(self.source_size, self.destination_size, float(self.active_length),
                            float(self.shift), self.kernel, self.taps, float(self.b),
                            float(self.c), float(self.blur), self.border, self.f64_mode)


@dataclass(frozen=True)
class IntegerConversion:
    """dsmvc::IntegerConversion (reference include/dsmvc/engine.hpp:167-173)."""
    input_offset: float
    input_scale: float
    output_scale: float
    output_offset: float
    output_maximum: int

    def to_c(self) -> ConversionC:
        return ConversionC(self.input_offset, self.input_scale, self.output_scale,
                           self.output_offset, self.output_maximum)


def integer_conversion(bits: int, chroma: bool, full_range: bool) -> IntegerConversion:
    """Per-plane integer constants of a YUV clip (reference src/vs_plugin.cpp:761-781).
    `full_range` follows the reference's frame_range(): _Range / _ColorRange == 1
    means full (vs_plugin.cpp:750-759)."""
    maximum = (1 << bits) - 1
    offset, scale = 0, maximum
    if not full_range:
        depth = 1 << (bits - 8)
        offset = (128 if chroma else 16) * depth
        scale = (224 if chroma else 219) * depth
    elif chroma:
        offset = 1 << (bits - 1)
    return IntegerConversion(float(offset), float(np.float32(1.0) / np.float32(scale)),
                             float(scale), float(offset), maximum)


class Plan:
    """Immutable axis plan built by the library's FP64 planner (cached)."""

    def __init__(self, request: AxisRequest, _handle: Optional[ctypes.c_void_p] = None):
        lib = load()
        if _handle is None:
            handle = ctypes.c_void_p()
            check(lib.dsmvc_hip_plan_build(ctypes.byref(request.to_c()), ctypes.byref(handle)))
        else:
            handle = _handle   # already built (build_plans), owned from here on
        self._handle = handle
        self.request = request
        info = PlanInfoC()
        check(lib.dsmvc_hip_plan_info_get(handle, ctypes.byref(info)))
        self.source_size = info.source_size
        self.destination_size = info.destination_size
        self.support = info.support
        self.half_bandwidth = info.half_bandwidth
        self.active_length = info.active_length
        self.shift = info.shift
        self.normal_rcond = info.normal_rcond
        self.normal_inf_norm = info.normal_inf_norm
        self.requires_float64 = bool(info.requires_float64)
        self.nnz = info.nnz

    @property
    def handle(self):
        return self._handle

    def __del__(self):
        try:
            if self._handle:
                load().dsmvc_hip_plan_release(self._handle)
        except Exception:
            pass

    def window_info(self) -> dict:
        """How the windowed (independent-chunk) kernels partition this plan."""
        info = _lib.WindowInfoC()
        check(load().dsmvc_hip_plan_window_info(self._handle, ctypes.byref(info)))
        return {name: getattr(info, name) for name, _ in _lib.WindowInfoC._fields_}

    def operator(self) -> dict:
        """The plan's explicit operator M = G^-1 A^T truncated to its band (the
        fused 2D route's operand): info fields plus `lo`, `count` and the
        concatenated row `values`."""
        info = _lib.OperatorInfoC()
        check(load().dsmvc_hip_plan_operator(self._handle, ctypes.byref(info), None, None, None, 0))
        n = self.destination_size
        lo, count = np.zeros(n, np.int32), np.zeros(n, np.int32)
        values = np.zeros(max(1, info.values))
        check(load().dsmvc_hip_plan_operator(self._handle, ctypes.byref(info),
                                             ctypes.c_void_p(lo.ctypes.data),
                                             ctypes.c_void_p(count.ctypes.data),
                                             ctypes.c_void_p(values.ctypes.data), values.size))
        out = {name: getattr(info, name) for name, _ in _lib.OperatorInfoC._fields_}
        out.update(lo=lo, count=count, values=values[:info.values])
        return out

    def arrays(self) -> dict:
        """Copies of the AxisPlan arrays (reference include/dsmvc/engine.hpp:90-107)."""
        n, w, nnz, f64 = self.destination_size, self.half_bandwidth, self.nnz, self.requires_float64
        out = {
            "offsets": np.zeros(n + 1, np.uint32), "indices": np.zeros(nnz, np.int32),
            "weights": np.zeros(nnz, np.float32), "lower_ld": np.zeros(w * n, np.float32),
            "upper_l": np.zeros(w * n, np.float32), "inverse_diagonal": np.zeros(n, np.float32),
            "weights_f64": np.zeros(nnz if f64 else 0), "normal_bands_f64":
                np.zeros((w + 1) * n if f64 else 0), "ldlt_bands_f64":
                np.zeros((w + 1) * n if f64 else 0), "inverse_diagonal_f64":
                np.zeros(n if f64 else 0),
        }
        ptrs = [ctypes.c_void_p(a.ctypes.data) if a.size else None for a in out.values()]
        check(load().dsmvc_hip_plan_arrays(self._handle, *ptrs))
        return out


CUSTOM = 6


def build_custom_plan(request: AxisRequest, kernel) -> Plan:
    """A plan for a custom kernel (dsmvc_hip_plan_build_custom): kernel(x) is
    called on the host at each tap's |distance| (blur applied) while the plan
    is built; support = request.taps.  Custom plans are never cached
    (reference src/axis_plan.cpp:535-599, 1367-1372)."""
    failure = []

    def call(x, _user):
        try:
            return float(kernel(x))
        except Exception as exc:   # reported as a non-finite value below
            failure.append(exc)
            return float("nan")

    fn = _lib.KERNEL_FN(call)
    req = AxisRequest(request.source_size, request.destination_size, request.active_length,
                      request.shift, CUSTOM, request.taps, request.b, request.c, request.blur,
                      request.border, request.f64_mode)
    handle = ctypes.c_void_p()
    status = load().dsmvc_hip_plan_build_custom(ctypes.byref(req.to_c()), fn, None,
                                                ctypes.byref(handle))
    if failure:
        raise failure[0]
    check(status)
    return Plan(req, _handle=handle)


def import_plan(fields, arrays) -> Plan:
    """A plan made by another planner -- for example the reference engine's
    AxisPlan -- from its scalar fields and arrays (dsmvc_hip_plan_import; the
    arrays are copied and checked by the reference's AxisPlan::valid() rules).
    `fields`: an object or dict with source_size, destination_size, support,
    half_bandwidth, active_length, shift, normal_rcond, normal_inf_norm and
    requires_float64; `arrays`: the AxisPlan arrays by name (Plan.arrays()
    names), the four FP64 arrays only when requires_float64."""
    get = (lambda k: fields[k]) if isinstance(fields, dict) else (lambda k: getattr(fields, k))
    f64 = bool(get("requires_float64"))
    kinds = {"offsets": np.uint32, "indices": np.int32, "weights_f64": np.float64,
             "normal_bands_f64": np.float64, "ldlt_bands_f64": np.float64,
             "inverse_diagonal_f64": np.float64}
    keep = {}
    for name in _lib.PLAN_ARRAYS:
        if name.endswith("_f64") and not f64:
            continue
        keep[name] = np.ascontiguousarray(arrays[name], dtype=kinds.get(name, np.float32))
    info = PlanInfoC(get("source_size"), get("destination_size"), get("support"),
                     get("half_bandwidth"), float(get("active_length")), float(get("shift")),
                     float(get("normal_rcond")), float(get("normal_inf_norm")), int(f64),
                     int(keep["indices"].size))
    data = _lib.PlanDataC(*[ctypes.c_void_p(keep[n].ctypes.data) if n in keep and keep[n].size
                            else None for n in _lib.PLAN_ARRAYS])
    handle = ctypes.c_void_p()
    check(load().dsmvc_hip_plan_import(ctypes.byref(info), ctypes.byref(data),
                                       ctypes.byref(handle)))
    return Plan(None, _handle=handle)


def build_plan(request: AxisRequest) -> Plan:
    return Plan(request)


def build_plans(requests: Sequence[AxisRequest], threads: int = 0) -> List[Plan]:
    """Many plans through the planner cache on host threads
    (dsmvc_hip_plan_build_many; threads <= 0: one per hardware thread, at most
    16).  Each plan equals Plan(request); a failure raises the error of the
    lowest failing request and returns nothing."""
    reqs = list(requests)
    if not reqs:
        return []
    array = (AxisRequestC * len(reqs))(*[r.to_c() for r in reqs])
    handles = (ctypes.c_void_p * len(reqs))()
    check(load().dsmvc_hip_plan_build_many(array, len(reqs), int(threads), handles))
    return [Plan(r, _handle=ctypes.c_void_p(h)) for r, h in zip(reqs, handles)]


def planner_cache_stats() -> dict:
    s = _lib.PlannerStatsC()
    check(load().dsmvc_hip_planner_cache_stats(ctypes.byref(s)))
    return {name: getattr(s, name) for name, _ in _lib.PlannerStatsC._fields_}


def clear_planner_caches():
    load().dsmvc_hip_clear_planner_caches()


def padding_index(index: int, size: int, mode: int) -> int:
    out = ctypes.c_int32()
    check(load().dsmvc_hip_padding_index(index, size, mode, ctypes.byref(out)))
    return out.value


def parse_backend(name: str) -> str:
    kind = ctypes.c_int32()
    check(load().dsmvc_hip_parse_backend(name.encode(), ctypes.byref(kind)))
    return BACKENDS[kind.value]


def resolve_backend(name: str) -> str:
    """Explicit `hip` without a device raises; it never falls back
    (reference src/backend.cpp:119-142)."""
    lib = load()
    kind = ctypes.c_int32()
    check(lib.dsmvc_hip_parse_backend(name.encode(), ctypes.byref(kind)))
    resolved = ctypes.c_int32()
    check(lib.dsmvc_hip_resolve_backend(kind.value, ctypes.byref(resolved)))
    return BACKENDS[resolved.value]


def device_count() -> int:
    out = ctypes.c_int32()
    check(load().dsmvc_hip_device_count(ctypes.byref(out)))
    return out.value


def _torch():
    import torch
    return torch


def _stream_handle(stream) -> ctypes.c_void_p:
    if stream is None:
        stream = _torch().cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _dev_ptr(t) -> ctypes.c_void_p:
    if not t.is_cuda:
        raise ValueError("dsmvc: HIP executor buffers must be device tensors")
    return ctypes.c_void_p(t.data_ptr())


def _sample_kind(t) -> Optional[int]:
    torch = _torch()
    return {torch.float32: 0, torch.uint8: 1, torch.uint16: 2}.get(t.dtype)


def _check_planes(t, what: str, rows: int, cols: int, dtype=None, frames: Optional[int] = None):
    """Checks, before any device call, that tensor `t` ([H, W] or [frames, H, W])
    holds at least `rows` x `cols` samples per plane, of `dtype` when given, and
    `frames` planes when given, with contiguous rows and strides that keep rows
    and planes apart -- so that no kernel reads or writes past the buffer.
    Returns t as [frames, H, W]."""
    if t.dim() not in (2, 3):
        raise ValueError(f"dsmvc: {what} must be a 2-D plane or a 3-D batch of planes")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"dsmvc: {what} must be {str(dtype).replace('torch.', '')}")
    tb = t if t.dim() == 3 else t.unsqueeze(0)
    n, h, w = tb.shape
    if frames is not None and n != frames:
        raise ValueError(f"dsmvc: {what} has {n} frames, expected {frames}")
    if h < rows or w < cols:
        raise ValueError(f"dsmvc: {what} is {h}x{w}, needs at least {rows}x{cols} samples")
    if tb.stride(2) != 1:
        raise ValueError(f"dsmvc: {what} rows must be contiguous")
    if (rows > 1 and tb.stride(1) < cols) or (n > 1 and tb.stride(0) < tb.stride(1) * rows):
        raise ValueError(f"dsmvc: {what} strides overlap its rows or planes")
    return tb


def _check_served(h: int, w: int, row_stride: int, col_stride: int, side: str, k: int,
                  rows: int, cols: int):
    """_check_planes for served plane k (input or output side) from its shape
    and strides."""
    if h < rows or w < cols:
        raise ValueError(f"dsmvc: {side} plane {k} is {h}x{w}, needs at least {rows}x{cols} "
                         f"samples")
    if col_stride != 1:
        raise ValueError(f"dsmvc: {side} plane {k} rows must be contiguous")
    if rows > 1 and row_stride < cols:
        raise ValueError(f"dsmvc: {side} plane {k} strides overlap its rows or planes")


class Executor:
    """dsmvc::Executor(backend="hip") on one GPU (reference include/dsmvc/engine.hpp:363-504).

    solve: "auto" (chunked partitioned solve where the plan allows, else
    sequential), "exact" (sequential, bit-identical to the ordered CPU solver),
    "chunked" (raise when not chunkable)."""

    def __init__(self, device: int = -1, solve: str = "auto"):
        lib = load()
        handle = ctypes.c_void_p()
        check(lib.dsmvc_hip_executor_create(device, SOLVE_MODES[solve], ctypes.byref(handle)))
        self._handle = handle
        self.solve = solve

    def __del__(self):
        try:
            if self._handle:
                load().dsmvc_hip_executor_destroy(self._handle)
                self._handle = None
        except Exception:
            pass

    def prepare(self, plan: Plan):
        check(load().dsmvc_hip_executor_prepare(self._handle, plan.handle))

    def seal(self):
        check(load().dsmvc_hip_executor_seal(self._handle))

    def tune(self, rows_chunks: int = 0, columns_chunks: int = 0):
        check(load().dsmvc_hip_executor_tune(self._handle, rows_chunks, columns_chunks))

    def tuned_route(self, plan: Plan, rows: bool, sample_kind: int = 0) -> int:
        """The autotuner's route for a plan and direction: a chunk count, -1
        (windowed kernels), or 0 when not tuned yet."""
        route = ctypes.c_int32()
        check(load().dsmvc_hip_executor_tuned_route(self._handle, plan.handle, int(rows),
                                                    sample_kind, ctypes.byref(route)))
        return route.value

    def set_refinement(self, steps: int):
        """Refinement steps per axis for executors in "refined" mode."""
        check(load().dsmvc_hip_executor_set_refinement(self._handle, steps))

    def synchronize(self, stream=None):
        check(load().dsmvc_hip_synchronize(self._handle, _stream_handle(stream)))

    # -- host frames (numpy arrays; staged through the executor's slots) -------
    def inverse_2d_host(self, horizontal: Plan, vertical: Plan, x, out=None,
                        conversion: Optional[IntegerConversion] = None):
        """x: host array [H, W] or [frames, H, W] of float32, uint8 or uint16.
        Returns a new (or fills `out`) host array [frames,] h x w."""
        kinds = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.uint16): 2}
        x = np.asarray(x)
        if x.dtype not in kinds:
            raise ValueError("dsmvc: samples must be float32, uint8 or uint16")
        xb = x if x.ndim == 3 else x[None]
        if xb.ndim != 3 or xb.shape[1] != vertical.source_size or xb.shape[2] < horizontal.source_size:
            raise ValueError("dsmvc: input plane does not match the plans")
        if xb.strides[2] != xb.itemsize:
            raise ValueError("dsmvc: rows must be contiguous")
        if kinds[x.dtype] and conversion is None:
            raise ValueError("dsmvc: integer planes need a conversion")
        shape = (xb.shape[0], vertical.destination_size, horizontal.destination_size)
        ob = np.empty(shape, x.dtype) if out is None else (out if out.ndim == 3 else out[None])
        if ob.shape != shape or ob.dtype != x.dtype or ob.strides[2] != ob.itemsize:
            raise ValueError("dsmvc: output does not match the plans")
        item = x.itemsize
        conv = ctypes.byref(conversion.to_c()) if conversion is not None else None
        check(load().dsmvc_hip_inverse_2d_host(
            self._handle, horizontal.handle, vertical.handle, kinds[x.dtype],
            ctypes.c_void_p(xb.ctypes.data), xb.strides[1] // item, xb.strides[0] // item,
            ctypes.c_void_p(ob.ctypes.data), ob.strides[1] // item, ob.strides[0] // item,
            xb.shape[0], conv))
        if out is not None:
            return out
        return ob if x.ndim == 3 else ob[0]

    def inverse_rows_host(self, plan: Plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] < plan.source_size:
            raise ValueError("dsmvc: input does not match the plan")
        out = np.empty((x.shape[0], plan.destination_size), np.float32)
        check(load().dsmvc_hip_inverse_rows_host(
            self._handle, plan.handle, ctypes.c_void_p(x.ctypes.data), x.shape[1],
            ctypes.c_void_p(out.ctypes.data), out.shape[1], x.shape[0]))
        return out

    def inverse_columns_host(self, plan: Plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[0] < plan.source_size:
            raise ValueError("dsmvc: input does not match the plan")
        out = np.empty((plan.destination_size, x.shape[1]), np.float32)
        check(load().dsmvc_hip_inverse_columns_host(
            self._handle, plan.handle, ctypes.c_void_p(x.ctypes.data), x.shape[1],
            ctypes.c_void_p(out.ctypes.data), out.shape[1], x.shape[1]))
        return out

    # -- device plan cache ------------------------------------------------------
    def inverse_2d_exported(self, horizontal: Plan, vertical: Plan, input_plane, output_plane,
                            sample_kind: int = 0, conversion: Optional[IntegerConversion] = None):
        """dsmvc_hip_inverse_2d_exported: planes of GPU frames exported by
        another API (`_lib.ExportedPlaneC`, e.g. a POSIX descriptor from
        VapourSynth's GPU core); both planes' handles are consumed."""
        conv = ctypes.byref(conversion.to_c()) if conversion is not None else None
        check(load().dsmvc_hip_inverse_2d_exported(
            self._handle, horizontal.handle, vertical.handle, sample_kind,
            ctypes.byref(input_plane), ctypes.byref(output_plane), conv))

    def inverse_rows_exported(self, plan: Plan, input_plane, output_plane, rows: int):
        check(load().dsmvc_hip_inverse_rows_exported(
            self._handle, plan.handle, ctypes.byref(input_plane), ctypes.byref(output_plane), rows))

    def inverse_columns_exported(self, plan: Plan, input_plane, output_plane, columns: int):
        check(load().dsmvc_hip_inverse_columns_exported(
            self._handle, plan.handle, ctypes.byref(input_plane), ctypes.byref(output_plane),
            columns))

    def import_exported_semaphore(self, semaphore: int, handle_type: int, handle: int) -> bool:
        out = ctypes.c_int32()
        check(load().dsmvc_hip_import_exported_semaphore(self._handle, semaphore, handle_type,
                                                         handle, ctypes.byref(out)))
        return bool(out.value)

    def has_exported_semaphore(self, semaphore: int) -> bool:
        out = ctypes.c_int32()
        check(load().dsmvc_hip_has_exported_semaphore(self._handle, semaphore, ctypes.byref(out)))
        return bool(out.value)

    def import_stats(self) -> dict:
        s = _lib.ImportStatsC()
        check(load().dsmvc_hip_executor_import_stats(self._handle, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.ImportStatsC._fields_}

    def set_plan_cache(self, limit_bytes: int):
        check(load().dsmvc_hip_executor_set_plan_cache(self._handle, int(limit_bytes)))

    def plan_cache_stats(self) -> dict:
        st = _lib.PlanCacheStatsC()
        check(load().dsmvc_hip_executor_plan_cache_stats(self._handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _lib.PlanCacheStatsC._fields_}

    # -- single axis, float32 ------------------------------------------------
    def inverse_rows(self, plan: Plan, x, out=None, stream=None, asynchronous=False):
        torch = _torch()
        if x.dim() != 2:
            raise ValueError("dsmvc: inverse_rows takes a 2-D plane")
        rows = x.shape[0]
        _check_planes(x, "input", rows, plan.source_size, torch.float32)
        if out is None:
            out = torch.empty((rows, plan.destination_size), dtype=torch.float32, device=x.device)
        if out.dim() != 2:
            raise ValueError("dsmvc: inverse_rows writes a 2-D plane")
        _check_planes(out, "output", rows, plan.destination_size, torch.float32)
        check(load().dsmvc_hip_inverse_rows(
            self._handle, plan.handle, _dev_ptr(x), x.stride(0), _dev_ptr(out), out.stride(0),
            rows, _stream_handle(stream), ASYNC if asynchronous else 0))
        return out

    def inverse_columns(self, plan: Plan, x, out=None, stream=None, asynchronous=False):
        torch = _torch()
        if x.dim() != 2:
            raise ValueError("dsmvc: inverse_columns takes a 2-D plane")
        cols = x.shape[1]
        if x.shape[0] < plan.source_size:
            raise ValueError("dsmvc: input has fewer rows than the plan's source size")
        _check_planes(x, "input", plan.source_size, cols, torch.float32)
        if out is None:
            out = torch.empty((plan.destination_size, cols), dtype=torch.float32, device=x.device)
        if out.dim() != 2:
            raise ValueError("dsmvc: inverse_columns writes a 2-D plane")
        _check_planes(out, "output", plan.destination_size, cols, torch.float32)
        check(load().dsmvc_hip_inverse_columns(
            self._handle, plan.handle, _dev_ptr(x), x.stride(0), _dev_ptr(out), out.stride(0),
            cols, _stream_handle(stream), ASYNC if asynchronous else 0))
        return out

    # -- 2D ------------------------------------------------------------------
    def inverse_2d(self, horizontal: Plan, vertical: Plan, x, out=None,
                   conversion: Optional[IntegerConversion] = None, stream=None,
                   asynchronous=False):
        """x: [H, W] or [frames, H, W]; float32, uint8 or uint16 (integer needs a
        conversion).  Returns [h, w] / [frames, h, w] of x's dtype."""
        torch = _torch()
        if x.dim() not in (2, 3):
            raise ValueError("dsmvc: input must be a 2-D plane or a 3-D batch of planes")
        batched = x.dim() == 3
        xb = x if batched else x.unsqueeze(0)
        frames, rows, cols = xb.shape
        if rows != vertical.source_size or cols < horizontal.source_size:
            raise ValueError("dsmvc: input plane does not match the plans")
        kind = _sample_kind(xb)
        if kind is None:
            raise ValueError("dsmvc: samples must be float32, uint8 or uint16")
        if kind != 0 and conversion is None:
            raise ValueError("dsmvc: integer planes need a conversion")
        _check_planes(xb, "input", vertical.source_size, horizontal.source_size)
        if out is None:
            out = torch.empty((frames, vertical.destination_size, horizontal.destination_size),
                              dtype=xb.dtype, device=xb.device)
        ob = _check_planes(out, "output", vertical.destination_size,
                           horizontal.destination_size, xb.dtype, frames)
        conv = ctypes.byref(conversion.to_c()) if conversion is not None else None
        check(load().dsmvc_hip_inverse_2d_batch(
            self._handle, horizontal.handle, vertical.handle, kind, _dev_ptr(xb), xb.stride(1),
            xb.stride(0), _dev_ptr(ob), ob.stride(1), ob.stride(0), frames, conv,
            _stream_handle(stream), ASYNC if asynchronous else 0))
        return out if batched or out.dim() == 2 else out[0]


    def inverse_2d_planes(self, jobs, stream=None, asynchronous=False):
        """Several planes in one set of launches (dsmvc_hip_inverse_2d_planes).
        jobs: (horizontal plan, vertical plan, x, conversion or None) with x
        [H, W] or [frames, H, W], all of one dtype.  Returns the outputs."""
        torch = _torch()
        lib = load()
        table = (_lib.PlaneJobC * max(1, len(jobs)))()
        outs, keep = [], []
        kind = None
        for k, (hp, vp, x, conversion) in enumerate(jobs):
            if x.dim() not in (2, 3):
                raise ValueError("dsmvc: input must be a 2-D plane or a 3-D batch of planes")
            xb = x if x.dim() == 3 else x.unsqueeze(0)
            frames, rows, cols = xb.shape
            if rows != vp.source_size or cols < hp.source_size or xb.stride(2) != 1:
                raise ValueError("dsmvc: input plane does not match the plans")
            k_ = _sample_kind(xb)
            if k_ is None or (kind is not None and k_ != kind):
                raise ValueError("dsmvc: planes must share one sample type")
            if k_ != 0 and conversion is None:
                raise ValueError("dsmvc: integer planes need a conversion")
            _check_planes(xb, "input", vp.source_size, hp.source_size)
            kind = k_
            out = torch.empty((frames, vp.destination_size, hp.destination_size), dtype=xb.dtype,
                              device=xb.device)
            conv = conversion.to_c() if conversion is not None else None
            keep.append(conv)
            table[k] = _lib.PlaneJobC(hp.handle, vp.handle, _dev_ptr(xb), xb.stride(1),
                                      xb.stride(0), _dev_ptr(out), out.stride(1), out.stride(0),
                                      frames, 0,
                                      ctypes.cast(ctypes.pointer(conv), ctypes.c_void_p)
                                      if conv is not None else None)
            outs.append(out if x.dim() == 3 else out[0])
        check(lib.dsmvc_hip_inverse_2d_planes(self._handle, table, len(jobs),
                                              0 if kind is None else kind,
                                              _stream_handle(stream),
                                              ASYNC if asynchronous else 0))
        return outs


class Batch:
    """Many 2D F32 descale jobs of uneven geometry solved together: one launch
    per (direction, band class, tap width) over device-resident job tables
    (include/dsmvc_hip.h, dsmvc_hip_batch_*).

    jobs: (horizontal plan, vertical plan, input, output) with input a
    [frames,] H x W float32 device tensor and output [frames,] h x w.
    graph: replay the table launches from a captured HIP graph."""

    def __init__(self, executor: "Executor", jobs, graph: bool = False):
        lib = load()
        self._executor = executor
        self._keep = []
        table = (_lib.BatchJobC * max(1, len(jobs)))()
        torch = _torch()
        for k, (hp, vp, x, out) in enumerate(jobs):
            frames = x.shape[0] if x.dim() == 3 else 1
            if out.shape[-2:] != (vp.destination_size, hp.destination_size) or \
                    x.shape[-2] < vp.source_size or x.shape[-1] < hp.source_size:
                raise ValueError("dsmvc: batch job shapes do not match its plans")
            _check_planes(x, "batch input", vp.source_size, hp.source_size, torch.float32)
            _check_planes(out, "batch output", vp.destination_size, hp.destination_size,
                          torch.float32, frames)
            table[k] = _lib.BatchJobC(
                hp.handle, vp.handle, _dev_ptr(x), x.stride(-2),
                x.stride(0) if x.dim() == 3 else 0, _dev_ptr(out), out.stride(-2),
                out.stride(0) if out.dim() == 3 else 0, frames, 0)
            self._keep += [hp, vp, x, out]
        handle = ctypes.c_void_p()
        check(lib.dsmvc_hip_batch_create(executor._handle, table, len(jobs),
                                         _lib.BATCH_GRAPH if graph else 0, ctypes.byref(handle)))
        self._handle = handle

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            load().dsmvc_hip_batch_destroy(handle)
            self._handle = None

    def run(self, stream=None, asynchronous=False):
        check(load().dsmvc_hip_batch_run(self._handle, _stream_handle(stream),
                                         ASYNC if asynchronous else 0))

    def info(self) -> dict:
        info = _lib.BatchInfoC()
        check(load().dsmvc_hip_batch_info_get(self._handle, ctypes.byref(info)))
        return {name: getattr(info, name) for name, _ in _lib.BatchInfoC._fields_
                if name != "reserved"}


class Server:
    """Continuous-batching request loop (include/dsmvc_hip.h,
    dsmvc_hip_server_*): requests of any geometry are planned on submit,
    admitted as slots and device memory free up, and solved together with
    whatever else is admitted in the same round.  Planes are float32 device
    tensors that must stay alive until the request finishes (the Server keeps
    references until `wait`)."""

    FUSED = {"warm": 0, "always": 1, "off": 2}

    def __init__(self, executor: "Executor", memory_budget: int = 1 << 30, slots: int = 4,
                 max_round_requests: int = 256, fused: str = "warm"):
        """`fused`: when F32 planes take the fused 2D route -- "warm" once
        their plan pair's operator images exist (built off the submit path),
        "always" (built at submit; the route never changes), "off"."""
        self._executor = executor
        if fused not in self.FUSED:
            raise ValueError(f"dsmvc: fused must be one of {sorted(self.FUSED)}")
        cfg = _lib.ServerConfigC(memory_budget, slots, max_round_requests, self.FUSED[fused], 0)
        handle = ctypes.c_void_p()
        check(load().dsmvc_hip_server_create(executor._handle, ctypes.byref(cfg),
                                             ctypes.byref(handle)))
        self._handle = handle
        self._keep = {}
        self._local = threading.local()
        self._submit = load().dsmvc_hip_server_submit
        self._rid_local = threading.local()

    def close(self):
        handle = getattr(self, "_handle", None)
        if handle:
            load().dsmvc_hip_server_destroy(handle)
            self._handle = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def submit(self, planes_in, planes_out, kernel: int = BICUBIC, taps: int = 3, b: float = 0.0,
               c: float = 0.5, blur: float = 1.0, border: int = SYMMETRIC, f64_mode: int = 0,
               subsampling_w: int = 1, subsampling_h: int = 1) -> int:
        """planes_in / planes_out: 1 (gray) or 3 (YUV) 2D float32 tensors.

        Every plane is checked before the library sees it (float32, 2-D,
        on the device, contiguous rows, at least the size the library
        derives from luma and the subsampling); the request record is
        reused per calling thread, since the library copies it."""
        n = len(planes_in)
        if n not in (1, 3) or len(planes_out) != n:
            raise ValueError("dsmvc: a request has 1 or 3 planes")
        r = getattr(self._local, "request", None)
        if r is None:
            r = self._local.request = _lib.ServeRequestC()
        f32 = _torch().float32
        sub_w = subsampling_w if n == 3 else 0
        sub_h = subsampling_h if n == 3 else 0
        for k in range(n):
            x, o = planes_in[k], planes_out[k]
            if x.dtype != f32 or o.dtype != f32:
                raise ValueError("dsmvc: served planes are float32")
            if x.dim() != 2 or o.dim() != 2:
                raise ValueError("dsmvc: served planes are 2-D")
            (xh, xw), (xs, xc) = x.shape, x.stride()
            (oh, ow), (os_, oc) = o.shape, o.stride()
            if k == 0:
                src_h, src_w, dst_h, dst_w = xh, xw, oh, ow
            # The library derives chroma sizes from luma by the subsampling.
            sw_, sh_ = (sub_w, sub_h) if k else (0, 0)
            _check_served(xh, xw, xs, xc, "input", k, src_h >> sh_, src_w >> sw_)
            _check_served(oh, ow, os_, oc, "output", k, dst_h >> sh_, dst_w >> sw_)
            r.src_stride[k] = xs
            r.dst_stride[k] = os_
        for k in range(n):
            x, o = planes_in[k], planes_out[k]
            if not x.is_cuda or not o.is_cuda:
                raise ValueError("dsmvc: HIP executor buffers must be device tensors")
            r.src[k] = x.data_ptr()
            r.dst[k] = o.data_ptr()
        r.planes = n
        r.subsampling_w = sub_w
        r.subsampling_h = sub_h
        r.src_height, r.src_width, r.dst_height, r.dst_width = src_h, src_w, dst_h, dst_w
        r.kernel, r.taps, r.border, r.f64_mode = kernel, taps, border, f64_mode
        r.b, r.c, r.blur = b, c, blur
        rid = getattr(self._rid_local, "rid", None)
        if rid is None:
            rid = self._rid_local.rid = ctypes.c_uint64()
        check(self._submit(self._handle, ctypes.byref(r), ctypes.byref(rid)))
        self._keep[rid.value] = (planes_in, planes_out)
        return rid.value

    def wait(self, rid: int):
        """Blocks until the request finishes; returns (submitted, completed)
        in seconds since the server started.  Raises on a failed request."""
        t0, t1 = ctypes.c_double(), ctypes.c_double()
        lib = load()
        status = lib.dsmvc_hip_server_wait(self._handle, rid, ctypes.byref(t0), ctypes.byref(t1))
        self._keep.pop(rid, None)
        if status != _lib.OK:
            message = lib.dsmvc_hip_last_error().decode()
            lib.dsmvc_hip_server_release(self._handle, rid)
            raise (ValueError if status == _lib.INVALID_ARGUMENT else DsmvcError)(message)
        check(lib.dsmvc_hip_server_release(self._handle, rid))
        return t0.value, t1.value

    def stats(self) -> dict:
        s = _lib.ServerStatsC()
        check(load().dsmvc_hip_server_stats_get(self._handle, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.ServerStatsC._fields_}


# ---------------------------------------------------------------------------
# Plugin-shaped operators
# ---------------------------------------------------------------------------

@dataclass
class Clip:
    """A clip: planes (luma first) of [frames, H, W] device tensors plus format.

    color_family: "gray" or "yuv"; subsampling_w/h: log2 chroma subsampling;
    bits: bits per integer sample (ignored for float32); full_range: the
    _Range / _ColorRange frame property (True when it is 1)."""
    planes: List
    color_family: str = "gray"
    subsampling_w: int = 0
    subsampling_h: int = 0
    bits: int = 0
    full_range: bool = False

    @property
    def width(self):
        return self.planes[0].shape[-1]

    @property
    def height(self):
        return self.planes[0].shape[-2]


_executors = {}


def _executor(solve: str = "auto"):
    if solve not in _executors:
        _executors[solve] = Executor(-1, solve)
    return _executors[solve]


def filter_requests(source_width: int, source_height: int, width: int, height: int,
                    num_planes: int = 1, subsampling_w: int = 0, subsampling_h: int = 0,
                    src_left: float = 0.0, src_top: float = 0.0,
                    src_width: Optional[float] = None, src_height: Optional[float] = None,
                    kernel: int = BICUBIC, taps: int = 3, b: float = 0.0, c: float = 0.5,
                    blur: float = 1.0, border: int = SYMMETRIC, f64_mode: int = 0):
    """Axis requests per plane kind, as the reference's prepare_filter_requests
    derives them (src/vs_plugin.cpp:656-705): ([luma, chroma] horizontal,
    [luma, chroma] vertical), chroma present only for subsampled multi-plane
    clips.  Horizontal chroma is left-sited (shift 0.25 - 0.25 * width /
    source_width, plus src_left scaled to the chroma plane); vertical chroma
    takes src_top scaled only.  The active lengths are src_width / src_height
    (defaulting to the output size) scaled to each plane."""
    src_width = float(width) if src_width is None else float(src_width)
    src_height = float(height) if src_height is None else float(src_height)
    kw = dict(kernel=kernel, taps=taps, b=b, c=c, blur=blur, border=border, f64_mode=f64_mode)
    hreq = [AxisRequest(source_width, width, src_width, src_left, **kw)]
    vreq = [AxisRequest(source_height, height, src_height, src_top, **kw)]
    if num_planes > 1 and subsampling_w > 0:
        cw = source_width >> subsampling_w
        hreq.append(AxisRequest(cw, width >> subsampling_w,
                                src_width * float(cw) / float(source_width),
                                0.25 - 0.25 * float(width) / float(source_width)
                                + src_left * float(cw) / float(source_width), **kw))
    if num_planes > 1 and subsampling_h > 0:
        ch = source_height >> subsampling_h
        vreq.append(AxisRequest(ch, height >> subsampling_h,
                                src_height * float(ch) / float(source_height),
                                src_top * float(ch) / float(source_height), **kw))
    return hreq, vreq


def _to_float(plane, conv: IntegerConversion):
    """Integer samples to normalised float32, the same F32 operations as the
    fused kernels' load: (float(x) - offset) * scale."""
    torch = _torch()
    off = torch.tensor(conv.input_offset, dtype=torch.float32, device=plane.device)
    scale = torch.tensor(conv.input_scale, dtype=torch.float32, device=plane.device)
    return (plane.to(torch.float32) - off) * scale


def _to_integer(plane, conv: IntegerConversion, dtype):
    """float32 to integer samples: nearbyint(clamp(v * scale + offset, 0, max))
    under round-half-even, as the reference's integer output conversion."""
    torch = _torch()
    scale = torch.tensor(conv.output_scale, dtype=torch.float32, device=plane.device)
    off = torch.tensor(conv.output_offset, dtype=torch.float32, device=plane.device)
    v = (plane * scale + off).clamp(0.0, float(conv.output_maximum))
    return torch.round(v).to(torch.int32).to(dtype)


def Descale(src, width: int, height: int, kernel: Optional[str] = None,
            taps: Optional[int] = None, b: float = 0.0, c: float = 0.5, src_left: float = 0.0,
            src_top: float = 0.0, src_width: Optional[float] = None,
            src_height: Optional[float] = None, border_handling: Optional[int] = None,
            padding: Optional[int] = None, f64mode: int = 0, blur: float = 1.0,
            force: bool = False, force_h: Optional[bool] = None, force_v: Optional[bool] = None,
            custom=None, support: Optional[int] = None, custom_kernel=None,
            backend: str = "hip", solve: str = "auto"):
    """core.dsmvc.Descale with backend="hip" (reference src/vs_plugin.cpp:362-509,
    656-705, 1774-1987).  `src` is a Clip or a single float32 plane tensor.
    Integer clips descaled on both axes take the fused integer route; on one
    axis they are converted to float32, descaled and converted back, as the
    reference does through its resize calls (convert_and_invoke,
    vs_plugin.cpp:570-626).  A custom kernel (`custom_kernel`, or the legacy
    `custom`, a callable of the distance) needs `taps` or `support` and is
    planned on the host through it (vs_plugin.cpp:430-486, 628-653)."""
    if resolve_backend(backend) != "hip":
        raise ValueError("dsmvc: this build provides backend=\"hip\" only")
    clip = src if isinstance(src, Clip) else Clip([src])
    custom_fn = custom if custom is not None else custom_kernel
    if kernel is not None and custom_fn is not None:
        raise ValueError("specify either kernel or custom kernel, not both")
    if kernel is None and custom_fn is None:
        raise ValueError("kernel or custom kernel is required")
    if custom_fn is None and kernel not in KERNELS:
        raise ValueError("kernel is invalid")
    taps_given = taps is not None or support is not None
    taps = taps if taps is not None else (support if support is not None else 3)
    if (kernel == "lanczos" or custom_fn is not None) and taps < 1:
        raise ValueError("taps must be greater than zero")
    if custom_fn is not None and not taps_given:
        raise ValueError("custom kernels require taps or support")
    if width < 1:
        raise ValueError("width must be greater than zero")
    if height < 8:
        raise ValueError("height must be at least 8")
    if width > clip.width or height > clip.height:
        raise ValueError("output dimensions must not exceed input dimensions")
    if (src_width is not None and not src_width > 0.0) or \
            (src_height is not None and not src_height > 0.0):
        raise ValueError("src_width and src_height must be positive")
    if padding is not None and border_handling is not None:
        raise ValueError("specify either padding or border_handling, not both")
    if padding is not None:
        if padding not in (0, 1, 2, 3):
            raise ValueError("padding must be 0, 1, 2, or 3")
        border = padding
    elif border_handling is not None:
        border = {1: ZERO, 2: REPEAT}.get(border_handling, MIRROR)
    else:
        border = SYMMETRIC
    if f64mode not in (0, 1, 2):
        raise ValueError("f64mode must be 0, 1, or 2")
    if not (blur > 0.0) or not math.isfinite(blur):
        raise ValueError("blur must be finite and greater than zero")
    sw, sh = clip.width, clip.height
    extent = min(sw, sh)
    if len(clip.planes) > 1:
        extent = min(extent, sw >> clip.subsampling_w, sh >> clip.subsampling_h)
    if blur >= float(extent):
        raise ValueError("blur exceeds the supported source-plane extent")
    if clip.color_family == "yuv" and clip.subsampling_w | clip.subsampling_h:
        if width % (1 << clip.subsampling_w) or height % (1 << clip.subsampling_h):
            raise ValueError("output width is incompatible with subsampling")
    sw_act = float(width) if src_width is None else float(src_width)
    sh_act = float(height) if src_height is None else float(src_height)
    force_h = force if force_h is None else force_h
    force_v = force if force_v is None else force_v
    process_h = width != sw or src_left != 0.0 or sw_act != float(width) or force_h
    process_v = height != sh or src_top != 0.0 or sh_act != float(height) or force_v
    if not process_h and not process_v:
        return clip if isinstance(src, Clip) else src
    hreq, vreq = filter_requests(sw, sh, width, height, len(clip.planes), clip.subsampling_w,
                                 clip.subsampling_h, src_left, src_top, sw_act, sh_act,
                                 kernel=CUSTOM if custom_fn is not None else KERNELS[kernel],
                                 taps=taps, b=b, c=c, blur=blur, border=border, f64_mode=f64mode)
    make_plan = (lambda r: build_custom_plan(r, custom_fn)) if custom_fn is not None else Plan
    ex = _executor(solve)
    torch = _torch()

    def conversion_of(index, plane):
        if plane.dtype == torch.float32:
            return None
        return integer_conversion(clip.bits or (8 if plane.dtype == torch.uint8 else 16),
                                  index > 0 and clip.color_family == "yuv", clip.full_range)

    def plans_of(index):
        chroma = index > 0
        return (make_plan(hreq[1 if chroma and len(hreq) > 1 else 0]),
                make_plan(vreq[1 if chroma and len(vreq) > 1 else 0]))

    if process_h and process_v:
        # Every plane of the frame batch in one set of launches.
        jobs = [plans_of(i) + (plane, conversion_of(i, plane))
                for i, plane in enumerate(clip.planes)]
        out_planes = ex.inverse_2d_planes(jobs)
    else:
        out_planes = []
        for index, plane in enumerate(clip.planes):
            hp, vp = plans_of(index)
            conv = conversion_of(index, plane)
            x = plane if plane.dim() == 3 else plane.unsqueeze(0)
            xf = _to_float(x, conv) if conv is not None else x
            if process_h:
                res = torch.stack([ex.inverse_rows(hp, f) for f in xf])
            else:
                res = torch.stack([ex.inverse_columns(vp, f) for f in xf])
            if conv is not None:
                res = _to_integer(res, conv, plane.dtype)
            out_planes.append(res if plane.dim() == 3 else res[0])
    if isinstance(src, Clip):
        return Clip(out_planes, clip.color_family, clip.subsampling_w, clip.subsampling_h,
                    clip.bits, clip.full_range)
    return out_planes[0]


def Debilinear(src, width, height, **kw):
    return Descale(src, width, height, kernel="bilinear", **kw)


def Debicubic(src, width, height, b=0.0, c=0.5, **kw):
    return Descale(src, width, height, kernel="bicubic", b=b, c=c, **kw)


def Delanczos(src, width, height, taps=3, **kw):
    return Descale(src, width, height, kernel="lanczos", taps=taps, **kw)


def Despline16(src, width, height, **kw):
    return Descale(src, width, height, kernel="spline16", **kw)


def Despline36(src, width, height, **kw):
    return Descale(src, width, height, kernel="spline36", **kw)


def Despline64(src, width, height, **kw):
    return Descale(src, width, height, kernel="spline64", **kw)

"""Test-side helpers: ctypes bindings for the oracle (oracle/liboracle.so) and
for the reference engine built from its own sources (oracle/_ref/libdsmvc_ref.so),
plus restatements of the reference's golden-hash machinery.

Reference files restated here (paths relative to the reference repository):
  StableHasher / FNV-1a 64           tests/numerical_conformance.hpp:24-79
  make_normal_input (xorshift32)     tests/numerical_conformance.hpp:178-191
  f32_plan_hash / complete_plan_hash tests/numerical_conformance.hpp:201-265
  axis fixtures and golden hashes    tests/numerical_conformance.hpp:126-164
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libdsmvc_ref.so"

# dsmvc::KernelKind / BorderMode / F64Mode (reference include/dsmvc/engine.hpp:16-58)
BILINEAR, BICUBIC, LANCZOS, SPLINE16, SPLINE36, SPLINE64 = range(6)
ZERO, REPEAT, REFLECT101, SYMMETRIC, MIRROR = range(5)
F64_AUTO, F64_NEVER, F64_ALWAYS = range(3)


# --------------------------------------------------------------------------
# ctypes structures shared by the oracle and the reference shim
# --------------------------------------------------------------------------
class Request(ctypes.Structure):
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


def make_request(source, destination, active_length=None, shift=0.0, kernel=BICUBIC,
                 taps=3, b=0.0, c=0.5, blur=1.0, border=SYMMETRIC, f64_mode=F64_AUTO):
    return Request(source, destination,
                   float(destination if active_length is None else active_length),
                   float(shift), kernel, taps, float(b), float(c), float(blur), border,
                   f64_mode)


class Conversion(ctypes.Structure):
    _fields_ = [
        ("input_offset", ctypes.c_float),
        ("input_scale", ctypes.c_float),
        ("output_scale", ctypes.c_float),
        ("output_offset", ctypes.c_float),
        ("output_maximum", ctypes.c_uint32),
    ]


class OrcPlan(ctypes.Structure):
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
        ("offsets", ctypes.POINTER(ctypes.c_uint32)),
        ("indices", ctypes.POINTER(ctypes.c_int32)),
        ("weights", ctypes.POINTER(ctypes.c_float)),
        ("lower_ld", ctypes.POINTER(ctypes.c_float)),
        ("upper_l", ctypes.POINTER(ctypes.c_float)),
        ("inverse_diagonal", ctypes.POINTER(ctypes.c_float)),
        ("weights_f64", ctypes.POINTER(ctypes.c_double)),
        ("normal_bands_f64", ctypes.POINTER(ctypes.c_double)),
        ("ldlt_bands_f64", ctypes.POINTER(ctypes.c_double)),
        ("inverse_diagonal_f64", ctypes.POINTER(ctypes.c_double)),
    ]


class RefPlanInfo(ctypes.Structure):
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


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class PlanArrays:
    """Host copy of one axis plan (the AxisPlan fields, reference engine.hpp:82-107)."""
    source_size: int
    destination_size: int
    support: int
    half_bandwidth: int
    active_length: float
    shift: float
    normal_rcond: float
    normal_inf_norm: float
    requires_float64: bool
    offsets: np.ndarray
    indices: np.ndarray
    weights: np.ndarray
    lower_ld: np.ndarray
    upper_l: np.ndarray
    inverse_diagonal: np.ndarray
    weights_f64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    normal_bands_f64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ldlt_bands_f64: np.ndarray = field(default_factory=lambda: np.zeros(0))
    inverse_diagonal_f64: np.ndarray = field(default_factory=lambda: np.zeros(0))


# --------------------------------------------------------------------------
# Oracle bindings
# --------------------------------------------------------------------------
class Oracle:
    def __init__(self, path: Path = ORACLE_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built; run make -C oracle")
        lib = ctypes.CDLL(str(path))
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_plan_build.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(OrcPlan)]
        lib.orc_plan_free.argtypes = [ctypes.POINTER(OrcPlan)]
        lib.orc_padding_index.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]
        lib.orc_padding_index.restype = ctypes.c_int32
        vp, pd = ctypes.c_void_p, ctypes.c_ssize_t
        for name in ("orc_axis_f32", "orc_axis_f32_ordered", "orc_axis_f64_ordered",
                     "orc_axis_f64"):
            getattr(lib, name).argtypes = [ctypes.POINTER(OrcPlan), vp, pd, vp, pd]
        for name in ("orc_rows", "orc_columns"):
            getattr(lib, name).argtypes = [ctypes.POINTER(OrcPlan), vp, pd, vp, pd,
                                           ctypes.c_int32]
        for name in ("orc_2d", "orc_2d_ordered", "orc_2d_f64_exact"):
            getattr(lib, name).argtypes = [ctypes.POINTER(OrcPlan), ctypes.POINTER(OrcPlan),
                                           vp, pd, vp, pd]
        for name in ("orc_2d_u8", "orc_2d_u16"):
            getattr(lib, name).argtypes = [ctypes.POINTER(OrcPlan), ctypes.POINTER(OrcPlan),
                                           vp, pd, vp, pd, ctypes.POINTER(Conversion)]
        self.lib = lib

    def error(self):
        return self.lib.orc_last_error().decode()

    def _check(self, status):
        if status != 0:
            raise (ValueError if status == 1 else RuntimeError)(self.error())

    def build(self, request: Request) -> "OraclePlan":
        plan = OrcPlan()
        self._check(self.lib.orc_plan_build(ctypes.byref(request), ctypes.byref(plan)))
        return OraclePlan(self, plan)

    def padding_index(self, index, size, mode):
        return self.lib.orc_padding_index(index, size, mode)

    def axis_f32(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(plan.c.destination_size, np.float32)
        self._check(self.lib.orc_axis_f32(plan.ref, _ptr(x), 1, _ptr(out), 1))
        return out

    def axis_f32_ordered(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(plan.c.destination_size, np.float32)
        self.lib.orc_axis_f32_ordered(plan.ref, _ptr(x), 1, _ptr(out), 1)
        return out

    def axis_f64_ordered(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(plan.c.destination_size, np.float64)
        self.lib.orc_axis_f64_ordered(plan.ref, _ptr(x), 1, _ptr(out), 1)
        return out

    def axis_f64(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(plan.c.destination_size, np.float64)
        self.lib.orc_axis_f64(plan.ref, _ptr(x), 1, _ptr(out), 1)
        return out

    def rows(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((x.shape[0], plan.c.destination_size), np.float32)
        self._check(self.lib.orc_rows(plan.ref, _ptr(x), x.shape[1], _ptr(out),
                                      out.shape[1], x.shape[0]))
        return out

    def columns(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((plan.c.destination_size, x.shape[1]), np.float32)
        self._check(self.lib.orc_columns(plan.ref, _ptr(x), x.shape[1], _ptr(out),
                                         out.shape[1], x.shape[1]))
        return out

    def inverse_2d(self, h, v, x, ordered=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((v.c.destination_size, h.c.destination_size), np.float32)
        fn = self.lib.orc_2d_ordered if ordered else self.lib.orc_2d
        self._check(fn(h.ref, v.ref, _ptr(x), x.shape[1], _ptr(out), out.shape[1]))
        return out

    def inverse_2d_f64_exact(self, h, v, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty((v.c.destination_size, h.c.destination_size), np.float64)
        self._check(self.lib.orc_2d_f64_exact(h.ref, v.ref, _ptr(x), x.shape[1],
                                              _ptr(out), out.shape[1]))
        return out

    def inverse_2d_int(self, h, v, x, conversion: Conversion):
        x = np.ascontiguousarray(x)
        out = np.empty((v.c.destination_size, h.c.destination_size), x.dtype)
        fn = self.lib.orc_2d_u8 if x.dtype == np.uint8 else self.lib.orc_2d_u16
        self._check(fn(h.ref, v.ref, _ptr(x), x.shape[1], _ptr(out), out.shape[1],
                       ctypes.byref(conversion)))
        return out


class OraclePlan:
    def __init__(self, oracle: Oracle, plan: OrcPlan):
        self.oracle = oracle
        self.c = plan
        self.ref = ctypes.byref(self.c)

    def __del__(self):
        try:
            self.oracle.lib.orc_plan_free(self.ref)
        except Exception:
            pass

    def arrays(self) -> PlanArrays:
        p = self.c
        n, w, nnz = p.destination_size, p.half_bandwidth, p.nnz

        def take(ptr, count, dtype):
            if count == 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        f64 = bool(p.requires_float64)
        return PlanArrays(
            p.source_size, n, p.support, w, p.active_length, p.shift, p.normal_rcond,
            p.normal_inf_norm, f64,
            take(p.offsets, n + 1, np.uint32), take(p.indices, nnz, np.int32),
            take(p.weights, nnz, np.float32), take(p.lower_ld, w * n, np.float32),
            take(p.upper_l, w * n, np.float32), take(p.inverse_diagonal, n, np.float32),
            take(p.weights_f64, nnz if f64 else 0, np.float64),
            take(p.normal_bands_f64, (w + 1) * n if f64 else 0, np.float64),
            take(p.ldlt_bands_f64, (w + 1) * n if f64 else 0, np.float64),
            take(p.inverse_diagonal_f64, n if f64 else 0, np.float64))


# --------------------------------------------------------------------------
# Reference engine (built from its own sources) bindings
# --------------------------------------------------------------------------
class Reference:
    CPU_AUTOMATIC, CPU_SCALAR, CPU_AVX2, CPU_NEON, CPU_AVX512 = range(5)

    def __init__(self, path: Path = REF_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built")
        lib = ctypes.CDLL(str(path))
        vp, pd = ctypes.c_void_p, ctypes.c_ssize_t
        lib.ref_last_error.restype = ctypes.c_char_p
        lib.ref_plan_build.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(vp)]
        lib.ref_plan_free.argtypes = [vp]
        self.kernel_fn = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)
        lib.ref_plan_build_custom.argtypes = [ctypes.POINTER(Request), self.kernel_fn, vp,
                                              ctypes.POINTER(vp)]
        lib.ref_plan_info_get.argtypes = [vp, ctypes.POINTER(RefPlanInfo)]
        lib.ref_plan_arrays.argtypes = [vp] + [vp] * 10
        for name in ("ref_inverse_axis_f32", "ref_inverse_axis_f32_ordered",
                     "ref_inverse_axis_f64_ordered", "ref_inverse_axis_f64"):
            getattr(lib, name).argtypes = [vp, vp, pd, vp, pd]
        lib.ref_executor_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
        lib.ref_executor_free.argtypes = [vp]
        lib.ref_executor_name.argtypes = [vp]
        lib.ref_executor_name.restype = ctypes.c_char_p
        for name in ("ref_inverse_rows", "ref_inverse_columns"):
            getattr(lib, name).argtypes = [vp, vp, vp, pd, vp, pd, ctypes.c_int32]
        lib.ref_inverse_2d.argtypes = [vp, vp, vp, vp, pd, vp, pd]
        for name in ("ref_inverse_2d_u8", "ref_inverse_2d_u16"):
            getattr(lib, name).argtypes = [vp, vp, vp, vp, pd, vp, pd,
                                           ctypes.POINTER(Conversion), ctypes.c_int32]
        self.lib = lib
        self._executors = {}

    def _check(self, status):
        if status != 0:
            msg = self.lib.ref_last_error().decode()
            raise (ValueError if status == 1 else RuntimeError)(msg)

    def build(self, request: Request) -> "RefPlan":
        handle = ctypes.c_void_p()
        self._check(self.lib.ref_plan_build(ctypes.byref(request), ctypes.byref(handle)))
        return RefPlan(self, handle)

    def build_custom(self, request: Request, fn) -> "RefPlan":
        """The reference planner's custom-kernel path; fn(x) at |distance|."""
        callback = self.kernel_fn(lambda x, _user: fn(x))
        handle = ctypes.c_void_p()
        self._check(self.lib.ref_plan_build_custom(ctypes.byref(request), callback, None,
                                                   ctypes.byref(handle)))
        return RefPlan(self, handle)

    def executor(self, path):
        if path not in self._executors:
            handle = ctypes.c_void_p()
            self._check(self.lib.ref_executor_create(path, ctypes.byref(handle)))
            self._executors[path] = handle
        return self._executors[path]

    def executor_name(self, path):
        return self.lib.ref_executor_name(self.executor(path)).decode()

    def axis_f32(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(plan.info.destination_size, np.float32)
        self._check(self.lib.ref_inverse_axis_f32(plan.handle, _ptr(x), 1, _ptr(out), 1))
        return out

    def axis_f32_ordered(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(plan.info.destination_size, np.float32)
        self.lib.ref_inverse_axis_f32_ordered(plan.handle, _ptr(x), 1, _ptr(out), 1)
        return out

    def axis_f64_ordered(self, plan, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(plan.info.destination_size, np.float64)
        self.lib.ref_inverse_axis_f64_ordered(plan.handle, _ptr(x), 1, _ptr(out), 1)
        return out

    def inverse_2d(self, h, v, x, path=CPU_SCALAR):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((v.info.destination_size, h.info.destination_size), np.float32)
        self._check(self.lib.ref_inverse_2d(self.executor(path), h.handle, v.handle,
                                            _ptr(x), x.shape[1], _ptr(out), out.shape[1]))
        return out

    def inverse_2d_int(self, h, v, x, conversion, path=CPU_SCALAR, streamed=False):
        x = np.ascontiguousarray(x)
        out = np.empty((v.info.destination_size, h.info.destination_size), x.dtype)
        fn = self.lib.ref_inverse_2d_u8 if x.dtype == np.uint8 else self.lib.ref_inverse_2d_u16
        self._check(fn(self.executor(path), h.handle, v.handle, _ptr(x), x.shape[1],
                       _ptr(out), out.shape[1], ctypes.byref(conversion), int(streamed)))
        return out


class RefPlan:
    def __init__(self, ref: Reference, handle):
        self.refl = ref
        self.handle = handle
        self.info = RefPlanInfo()
        ref.lib.ref_plan_info_get(handle, ctypes.byref(self.info))

    def __del__(self):
        try:
            self.refl.lib.ref_plan_free(self.handle)
        except Exception:
            pass

    def arrays(self) -> PlanArrays:
        i = self.info
        n, w, nnz = i.destination_size, i.half_bandwidth, i.nnz
        f64 = bool(i.requires_float64)
        a = PlanArrays(
            i.source_size, n, i.support, w, i.active_length, i.shift, i.normal_rcond,
            i.normal_inf_norm, f64, np.zeros(n + 1, np.uint32), np.zeros(nnz, np.int32),
            np.zeros(nnz, np.float32), np.zeros(w * n, np.float32),
            np.zeros(w * n, np.float32), np.zeros(n, np.float32),
            np.zeros(nnz if f64 else 0), np.zeros((w + 1) * n if f64 else 0),
            np.zeros((w + 1) * n if f64 else 0), np.zeros(n if f64 else 0))
        bufs = [a.offsets, a.indices, a.weights, a.lower_ld, a.upper_l, a.inverse_diagonal,
                a.weights_f64, a.normal_bands_f64, a.ldlt_bands_f64, a.inverse_diagonal_f64]
        self.refl.lib.ref_plan_arrays(self.handle,
                                      *[(_ptr(b) if b.size else None) for b in bufs])
        return a


def reference_available() -> bool:
    return REF_SO.exists()


# --------------------------------------------------------------------------
# Golden-hash machinery (reference tests/numerical_conformance.hpp)
# --------------------------------------------------------------------------
FNV_OFFSET = 1469598103934665603
FNV_PRIME = 1099511628211
MASK64 = (1 << 64) - 1


class StableHasher:
    """FNV-1a 64 over little-endian bytes (numerical_conformance.hpp:24-79)."""

    def __init__(self):
        self.value = FNV_OFFSET

    def _bytes(self, data: bytes):
        v = self.value
        for byte in data:
            v ^= byte
            v = (v * FNV_PRIME) & MASK64
        self.value = v

    def add_string(self, s: str):
        raw = s.encode()
        self.add_u(len(raw), 8)
        self._bytes(raw)

    def add_u(self, value: int, nbytes: int):
        self._bytes((value & ((1 << (8 * nbytes)) - 1)).to_bytes(nbytes, "little"))

    def add_i32(self, v):
        self.add_u(int(v) & 0xFFFFFFFF, 4)

    def add_u8(self, v):
        self.add_u(int(v), 1)

    def add_bool(self, v):
        self.add_u(1 if v else 0, 1)

    def add_f64(self, v):
        self._bytes(struct.pack("<d", float(v)))

    def add_span(self, arr: np.ndarray, dtype):
        arr = np.ascontiguousarray(arr, dtype=dtype)
        self.add_u(arr.size, 8)
        self._bytes(arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes())

    def hex(self):
        return f"{self.value:016x}"


def _hash_request(h: StableHasher, r: Request):
    h.add_i32(r.source_size)
    h.add_i32(r.destination_size)
    h.add_f64(r.active_length)
    h.add_f64(r.shift)
    h.add_u8(r.kernel)
    h.add_i32(r.taps)
    h.add_f64(r.b)
    h.add_f64(r.c)
    h.add_u8(r.border)
    h.add_u8(r.f64_mode)


def _hash_plan_head(h: StableHasher, p: PlanArrays):
    h.add_i32(p.source_size)
    h.add_i32(p.destination_size)
    h.add_i32(p.support)
    h.add_i32(p.half_bandwidth)
    h.add_f64(p.active_length)
    h.add_f64(p.shift)


def f32_plan_hash(r: Request, p: PlanArrays) -> str:
    h = StableHasher()
    h.add_string("dsmvc-f32-axis-plan-v1")
    _hash_request(h, r)
    _hash_plan_head(h, p)
    h.add_bool(p.normal_rcond < 1.0e-4)
    h.add_bool(p.requires_float64)
    h.add_span(p.offsets, np.uint32)
    h.add_span(p.indices, np.int32)
    h.add_span(p.weights, np.float32)
    h.add_span(p.lower_ld, np.float32)
    h.add_span(p.upper_l, np.float32)
    h.add_span(p.inverse_diagonal, np.float32)
    return h.hex()


def complete_plan_hash(r: Request, p: PlanArrays) -> str:
    h = StableHasher()
    h.add_string("dsmvc-axis-plan-v1")
    _hash_request(h, r)
    _hash_plan_head(h, p)
    h.add_f64(p.normal_rcond)
    h.add_f64(p.normal_inf_norm)
    h.add_span(p.offsets, np.uint32)
    h.add_span(p.indices, np.int32)
    h.add_span(p.weights, np.float32)
    h.add_span(p.lower_ld, np.float32)
    h.add_span(p.upper_l, np.float32)
    h.add_span(p.inverse_diagonal, np.float32)
    h.add_span(p.weights_f64, np.float64)
    h.add_span(p.normal_bands_f64, np.float64)
    h.add_span(p.ldlt_bands_f64, np.float64)
    h.add_span(p.inverse_diagonal_f64, np.float64)
    return h.hex()


def output_hash(values: np.ndarray, dtype=np.float32) -> str:
    h = StableHasher()
    h.add_string("dsmvc-ordered-output-v1")
    h.add_span(values, dtype)
    return h.hex()


def make_normal_input(count: int, seed: int) -> np.ndarray:
    """xorshift32 -> floats in +-[0.5, 1) (numerical_conformance.hpp:178-191)."""
    out = np.empty(count, np.uint32)
    state = seed & 0xFFFFFFFF
    for i in range(count):
        state ^= (state << 13) & 0xFFFFFFFF
        state ^= state >> 17
        state ^= (state << 5) & 0xFFFFFFFF
        bits = 0x3F000000 | (state & 0x007FFFFF)
        if state & 0x00800000:
            bits |= 0x80000000
        out[i] = bits
    return out.view(np.float32)


def ordered_float_bits(x: np.ndarray) -> np.ndarray:
    bits = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.int64)
    neg = (bits & 0x80000000) != 0
    return np.where(neg, (~bits) & 0xFFFFFFFF, bits | 0x80000000)


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.abs(ordered_float_bits(a) - ordered_float_bits(b))


@dataclass
class AxisFixture:
    name: str
    request: Request
    input_seed: int
    expected_half_bandwidth: int
    f32_plan_hash: str
    complete_plan_hash: str
    ordered_output_hash: str
    production_output_hash: str


def conditioned_lanczos2_request(mode):
    return make_request(1080, 980, 978.1, 0.95, LANCZOS, 2, border=SYMMETRIC, f64_mode=mode)


def axis_fixtures():
    """The reference's seven axis fixtures and their non-Apple goldens
    (tests/numerical_conformance.hpp:126-164)."""
    R = make_request
    return [
        AxisFixture("b1-symmetric-tail", R(41, 35, 35.0, 0.1875, BILINEAR, 3, border=SYMMETRIC,
                                           f64_mode=F64_NEVER),
                    0xB1001001, 1, "b7fa7ca018f14328", "61f3f18073fe9e11",
                    "71ea8a545feca1b7", "ea1b09290d8d2d47"),
        AxisFixture("b3-reflect101-tail", R(47, 39, 39.0, -0.3125, BICUBIC, 3,
                                            border=REFLECT101, f64_mode=F64_NEVER),
                    0xB3003003, 3, "9a194da63434e8e1", "a8ee51581e483134",
                    "c2ec387c8df33a7b", "7ec44696e0ae2e8c"),
        AxisFixture("b5-repeat-tail", R(59, 49, 49.0, 0.4375, LANCZOS, 3, border=REPEAT,
                                        f64_mode=F64_NEVER),
                    0xB5005005, 5, "368aafc55d47f98c", "d7d468d0dcc3798e",
                    "0230fd2270a8d447", "a3b5a9429aaac189"),
        AxisFixture("b7-zero-tail", R(67, 57, 57.0, -0.25, SPLINE64, 3, border=ZERO,
                                      f64_mode=F64_NEVER),
                    0xB7007007, 7, "68da8f096af1407f", "dd8b79d05deb000e",
                    "19b4db02b77aa1e6", "8333021572f444c1"),
        AxisFixture("conditioned-f32-control", conditioned_lanczos2_request(F64_NEVER),
                    0xF3209781, 3, "5bddb1a9d8c61225", "162056d561d75888",
                    "97b0409984eb7606", "90fc1f98834e937e"),
        AxisFixture("forced-f64-b3", R(73, 61, 60.5, 0.125, BICUBIC, 3, border=SYMMETRIC,
                                       f64_mode=F64_ALWAYS),
                    0xF6400002, 3, "8192da268d0b6655", "b9ea390d346a2c8e",
                    "d4b9c8ff994a2520", "d4b9c8ff994a2520"),
        AxisFixture("automatic-risk-f64", conditioned_lanczos2_request(F64_AUTO),
                    0xF6409781, 3, "88de56a7458f03fd", "b0ceea661114a07b",
                    "1c332683a9d93688", "1c332683a9d93688"),
    ]


MIXED_INPUT_SEED = 0x2DF64B3A
MIXED_OUTPUT_HASH = "a9dcffe6649bd96e"


def mixed_horizontal_request():
    return make_request(17, 13, 12.75, 0.125, BICUBIC, 3, border=SYMMETRIC, f64_mode=F64_NEVER)


# --------------------------------------------------------------------------
# The FP64 accuracy oracle and the Float32 bar
# --------------------------------------------------------------------------
def fp64_request(r: Request) -> Request:
    """The same request with f64mode = 2: its plan keeps the FP64 factors for
    any geometry, and the oracle's Float64 route on it is the FP64 accuracy
    oracle (reference src/axis_plan.cpp:909-940, 1193-1248)."""
    return make_request(r.source_size, r.destination_size, r.active_length, r.shift, r.kernel,
                        r.taps, r.b, r.c, r.blur, r.border, F64_ALWAYS)


F32_BAR = 2.0e-6


def f32_bar(reference_error: float) -> float:
    """Largest max-abs error against the FP64 oracle allowed for a Float32
    route: 2e-6, or twice the reference's own ordered Float32 path's error on
    the same input where that is larger (ill-conditioned plans, whose F32
    error no F32 solve can avoid)."""
    return max(F32_BAR, 2.0 * reference_error)

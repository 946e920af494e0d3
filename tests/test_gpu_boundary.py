"""GPU tests of the drop-in boundary beyond the device-pointer entries.

  * the reference-side binding (reference_binding/hip_executor.cpp, compiled
    against the reference's own header): plans from the reference's planner,
    host buffers, through the C ABI -- exact mode bit-identical to the
    ordered oracle, integer output within 1 code of the reference
    CpuExecutor's scalar path, automatic mode within the F32 bar;
  * dsmvc_hip_plan_import: plans made by the oracle solve bit-identically to
    the same plans built by the library;
  * host-frame entries equal the device entries bit for bit;
  * execution slots: synchronous calls from four threads overlap and each
    result equals its solo result;
  * the device plan cache stays within its byte limit and results survive
    eviction.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import (BICUBIC, LANCZOS, ROOT, SPLINE36, F64_NEVER, Conversion, Reference,
                     Request, f32_bar, fp64_request, make_request, reference_available,
                     ulp_distance)  # noqa: E402
from workloads import integer_conversion  # noqa: E402

DEV = torch.device("cuda:0")
BINDING_SO = ROOT / "oracle" / "_ref" / "libdsmvc_binding.so"


def _plan(r):
    return dh.Plan(dh.AxisRequest(r.source_size, r.destination_size, r.active_length, r.shift,
                                  r.kernel, r.taps, r.b, r.c, r.blur, r.border, r.f64_mode))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Binding:
    """ctypes over oracle/_ref/libdsmvc_binding.so (binding_shim.cpp)."""

    def __init__(self):
        lib = ctypes.CDLL(str(BINDING_SO))
        vp, pd, i32 = ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int32
        lib.bnd_last_error.restype = ctypes.c_char_p
        lib.bnd_create.argtypes = [i32, ctypes.POINTER(vp)]
        lib.bnd_free.argtypes = [vp]
        lib.bnd_prepare.argtypes = [vp, ctypes.POINTER(Request)]
        lib.bnd_inverse_2d.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(Request), i32,
                                       vp, pd, vp, pd, ctypes.POINTER(Conversion)]
        for name in ("bnd_inverse_rows", "bnd_inverse_columns"):
            getattr(lib, name).argtypes = [vp, ctypes.POINTER(Request), vp, pd, vp, pd, i32]
        self.lib = lib

    def check(self, status):
        if status:
            raise (ValueError if status == 1 else RuntimeError)(self.lib.bnd_last_error().decode())

    def executor(self, mode):
        h = ctypes.c_void_p()
        self.check(self.lib.bnd_create(mode, ctypes.byref(h)))
        return h

    def inverse_2d(self, ex, hr, vr, x, conv=None):
        kind = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.uint16): 2}[x.dtype]
        out = np.zeros((vr.destination_size, hr.destination_size), x.dtype)
        self.check(self.lib.bnd_inverse_2d(ex, ctypes.byref(hr), ctypes.byref(vr), kind, _ptr(x),
                                           x.shape[1], _ptr(out), out.shape[1],
                                           ctypes.byref(conv) if conv is not None else None))
        return out


@pytest.fixture(scope="module")
def binding():
    if not BINDING_SO.exists() or not reference_available():
        pytest.skip("oracle/_ref/libdsmvc_binding.so not built (reference sources absent)")
    return Binding()


def _identical(a, b):
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | ((a == 0) & (b == 0))))


def test_reference_binding_end_to_end(binding, oracle):
    """The reference engine's own plans and host frames through the binding."""
    ref = Reference()
    hr = make_request(1920, 1280, kernel=BICUBIC)
    vr = make_request(1080, 720, kernel=BICUBIC)
    rng = np.random.default_rng(40)
    x = rng.uniform(0, 1, (1080, 1920)).astype(np.float32)
    exact, auto = binding.executor(1), binding.executor(0)
    binding.check(binding.lib.bnd_prepare(exact, ctypes.byref(hr)))
    binding.check(binding.lib.bnd_prepare(exact, ctypes.byref(vr)))
    got = binding.inverse_2d(exact, hr, vr, x)
    ordered = oracle.inverse_2d(oracle.build(hr), oracle.build(vr), x, ordered=True)
    assert _identical(got, ordered)
    # Borrowed (unprepared) plans on the automatic executor: F32 bar.
    got = binding.inverse_2d(auto, hr, vr, x)
    ref64 = oracle.inverse_2d_f64_exact(oracle.build(fp64_request(hr)),
                                        oracle.build(fp64_request(vr)), x.astype(np.float64))
    err = float(np.abs(got.astype(np.float64) - ref64).max())
    e_ref = float(np.abs(ordered.astype(np.float64) - ref64).max())
    print(f"binding auto F32: vs FP64 {err:.3g} (reference F32 {e_ref:.3g})", flush=True)
    assert err <= f32_bar(e_ref)
    # u16 and u8 against the reference CpuExecutor's scalar path.
    for bits, dtype in ((16, np.uint16), (8, np.uint8)):
        xi = rng.integers(0, 1 << bits, (540, 960)).astype(dtype)
        hq, vq = make_request(960, 640, kernel=SPLINE36), make_request(540, 360, kernel=SPLINE36)
        conv = integer_conversion(bits, chroma=bits == 16, full_range=bits == 8)
        got = binding.inverse_2d(auto, hq, vq, xi, conv).astype(np.int32)
        expected = ref.inverse_2d_int(ref.build(hq), ref.build(vq), xi, conv,
                                      path=Reference.CPU_SCALAR).astype(np.int32)
        diff = int(np.abs(got - expected).max())
        print(f"binding u{bits}: max code diff vs reference CpuExecutor {diff}", flush=True)
        assert diff <= 1
    # Single-axis calls, exact: bit-identical to the ordered oracle.
    rows = rng.uniform(0, 1, (37, 1920)).astype(np.float32)
    out = np.zeros((37, 1280), np.float32)
    binding.check(binding.lib.bnd_inverse_rows(exact, ctypes.byref(hr), _ptr(rows), 1920,
                                               _ptr(out), 1280, 37))
    op = oracle.build(hr)
    assert _identical(out, np.stack([oracle.axis_f32_ordered(op, r) for r in rows]))
    cols = np.ascontiguousarray(rows.T)
    outc = np.zeros((1280, 37), np.float32)
    binding.check(binding.lib.bnd_inverse_columns(exact, ctypes.byref(hr), _ptr(cols), 37,
                                                  _ptr(outc), 37, 37))
    assert _identical(outc, np.stack([oracle.axis_f32_ordered(op, r) for r in rows]).T)
    # An invalid argument keeps the reference's exception kind.
    with pytest.raises(ValueError):
        binding.check(binding.lib.bnd_inverse_rows(exact, ctypes.byref(hr), _ptr(rows), 1919,
                                                   _ptr(out), 1280, 37))
    binding.lib.bnd_free(exact)
    binding.lib.bnd_free(auto)


def test_imported_oracle_plans_solve_like_built_plans(oracle):
    ex = dh.Executor(0, "exact")
    rng = np.random.default_rng(41)
    for hr, vr in ((make_request(960, 640, kernel=LANCZOS, taps=3, f64_mode=F64_NEVER),
                    make_request(540, 360, kernel=BICUBIC, f64_mode=F64_NEVER)),
                   (make_request(300, 250, kernel=LANCZOS, taps=4, blur=0.9, f64_mode=2),
                    make_request(160, 100, kernel=BICUBIC, f64_mode=2))):
        oh, ov = oracle.build(hr), oracle.build(vr)
        ah, av = oh.arrays(), ov.arrays()
        ih, iv = dh.import_plan(ah, ah.__dict__), dh.import_plan(av, av.__dict__)
        x = rng.uniform(0, 1, (vr.source_size, hr.source_size)).astype(np.float32)
        xt = torch.from_numpy(x).to(DEV)
        got = ex.inverse_2d(ih, iv, xt).cpu().numpy()
        built = ex.inverse_2d(_plan(hr), _plan(vr), xt).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), built.view(np.uint32))
        expected = oracle.inverse_2d(oh, ov, x, ordered=not ih.requires_float64)
        if ih.requires_float64:
            assert int(ulp_distance(got, expected).max()) <= 1
        else:
            assert _identical(got, expected)


def test_host_entries_equal_device_entries():
    ex = dh.Executor(0, "auto")
    rng = np.random.default_rng(42)
    hp, vp = _plan(make_request(1920, 1280)), _plan(make_request(1080, 720))
    x = rng.uniform(0, 1, (3, 1080, 1920)).astype(np.float32)
    dev = ex.inverse_2d(hp, vp, torch.from_numpy(x).to(DEV)).cpu().numpy()
    host = ex.inverse_2d_host(hp, vp, x)
    assert np.array_equal(dev, host)
    conv = dh.integer_conversion(16, False, False)
    xi = rng.integers(0, 65536, (2, 1080, 1920)).astype(np.uint16)
    xt = torch.from_numpy(xi.astype(np.int32)).to(DEV).to(torch.uint16)
    dev = ex.inverse_2d(hp, vp, xt, conversion=conv).to(torch.int32).cpu().numpy()
    host = ex.inverse_2d_host(hp, vp, xi, conversion=conv).astype(np.int32)
    assert np.array_equal(dev, host)
    rows = rng.uniform(0, 1, (100, 1920)).astype(np.float32)
    assert np.array_equal(ex.inverse_rows_host(hp, rows),
                          ex.inverse_rows(hp, torch.from_numpy(rows).to(DEV)).cpu().numpy())
    cols = rng.uniform(0, 1, (1080, 77)).astype(np.float32)
    assert np.array_equal(ex.inverse_columns_host(vp, cols),
                          ex.inverse_columns(vp, torch.from_numpy(cols).to(DEV)).cpu().numpy())


def test_synchronous_calls_from_threads_overlap():
    """Four threads each make synchronous 2D calls without a stream (the
    fmParallel pattern) on one executor: every result equals its solo result,
    and the calls overlap on the device -- with the executor's lock held
    through each call's wait they would run one after another.  The exact
    route's sequential kernels give each call milliseconds of low-occupancy
    device work, so the overlap is plain in wall-clock time."""
    ex = dh.Executor(0, "exact")
    rng = np.random.default_rng(43)
    hp, vp = _plan(make_request(1920, 1280)), _plan(make_request(1080, 720))
    xs = [torch.from_numpy(rng.uniform(0, 1, (1080, 1920)).astype(np.float32)).to(DEV)
          for _ in range(4)]
    torch.cuda.synchronize()
    solo = [ex.inverse_2d(hp, vp, x, stream=torch.cuda.default_stream()).clone() for x in xs]
    calls = 4

    def run(k, outs):
        for _ in range(calls):
            outs.append(ex.inverse_2d(hp, vp, xs[k], stream=torch.cuda.default_stream()))

    t0 = time.perf_counter()
    for k in range(4):
        run(k, [])
    serial = time.perf_counter() - t0
    results = [[] for _ in range(4)]
    threads = [threading.Thread(target=run, args=(k, results[k])) for k in range(4)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    parallel = time.perf_counter() - t0
    print(f"4 threads x {calls} synchronous exact 1080p calls: {parallel * 1e3:.1f} ms, one "
          f"after another {serial * 1e3:.1f} ms ({serial / parallel:.2f}x)", flush=True)
    for k in range(4):
        for out in results[k]:
            assert torch.equal(out, solo[k]), k
    assert parallel < 0.6 * serial


def test_plan_cache_stays_within_its_limit():
    ex = dh.Executor(0, "auto")
    limit = 6 << 20
    ex.set_plan_cache(limit)
    x = torch.rand(540, 960, device=DEV)
    first = None
    for k in range(40):
        hp = _plan(make_request(960, 500 + 8 * k))
        vp = _plan(make_request(540, 280 + 4 * k))
        out = ex.inverse_2d(hp, vp, x)
        if k == 0:
            first = (hp, vp, out.clone())
        stats = ex.plan_cache_stats()
        assert stats["bytes"] <= limit, stats
    print(ex.plan_cache_stats(), flush=True)
    assert ex.plan_cache_stats()["evictions"] > 0
    hp, vp, expected = first
    assert torch.equal(ex.inverse_2d(hp, vp, x), expected)

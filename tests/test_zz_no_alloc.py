"""dsmvc_hip_inverse_2d_batch_noalloc against dsmvc_hip_inverse_2d_batch.

The no-allocation entry runs the 2D route out of a fixed device array.  It is
checked through the C ABI against the allocating entry on the same inputs, on
two calls in a row (the second smaller, so stale contents of the shared array
would show), to the usual bars: Float32 max-abs <= 2e-6 (the windowed and
sequential kernels are the same ones, so results are in fact identical),
integer <= 1 code, Float64 route <= 1 ULP; sizes beyond the array, plans not
prepared, refined executors and asynchronous Float64 calls are refused.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from dsmvc_hip import _lib  # noqa: E402
from support import ulp_distance  # noqa: E402

DEV = torch.device("cuda:0")
KINDS = {torch.float32: 0, torch.uint8: 1, torch.uint16: 2}


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return _lib.load()


def _call(lib, entry, ex, hp, vp, x, out, conv=None, flags=0):
    conv_c = ctypes.byref(conv.to_c()) if conv is not None else None
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    status = getattr(lib, entry)(
        ex._handle, hp.handle, vp.handle, KINDS[x.dtype], ctypes.c_void_p(x.data_ptr()),
        x.stride(1), x.stride(0), ctypes.c_void_p(out.data_ptr()), out.stride(1), out.stride(0),
        x.shape[0], conv_c, stream, flags)
    return status, lib.dsmvc_hip_last_error().decode()


def _both(lib, ex, hp, vp, x, conv=None):
    h, w = vp.destination_size, hp.destination_size
    ref = torch.full((x.shape[0], h, w), 0, dtype=x.dtype, device=DEV)
    got = torch.full((x.shape[0], h, w), 0, dtype=x.dtype, device=DEV)
    assert _call(lib, "dsmvc_hip_inverse_2d_batch", ex, hp, vp, x, ref, conv)[0] == 0
    status, message = _call(lib, "dsmvc_hip_inverse_2d_batch_noalloc", ex, hp, vp, x, got, conv)
    assert status == 0, message
    return got, ref


def _plans(ex, sw, sh, dw, dh_, **kw):
    hp = dh.Plan(dh.AxisRequest(sw, dw, float(dw), **kw))
    vp = dh.Plan(dh.AxisRequest(sh, dh_, float(dh_), **kw))
    ex.prepare(hp)
    ex.prepare(vp)
    return hp, vp


def test_float32_then_smaller_integer_call(lib):
    ex = dh.Executor(0, "auto")
    rng = np.random.default_rng(3)
    hp, vp = _plans(ex, 1920, 1080, 1280, 720)
    x = torch.from_numpy(rng.uniform(0, 1, (2, 1080, 1920)).astype(np.float32)).to(DEV)
    got, ref = _both(lib, ex, hp, vp, x)
    err = float((got - ref).abs().max())
    print(f"F32 1080p->720p x2: max-abs vs allocating entry {err:.3g}", flush=True)
    assert err <= 2e-6
    # Second call, smaller: a u16 chroma-sized plane.
    hq, vq = _plans(ex, 960, 540, 640, 360, kernel=dh.LANCZOS, taps=3)
    conv = dh.integer_conversion(16, True, False)
    xi = torch.from_numpy(rng.integers(0, 65536, (1, 540, 960)).astype(np.int32)).to(DEV)
    xi = xi.to(torch.uint16)
    got, ref = _both(lib, ex, hq, vq, xi, conv)
    diff = int((got.to(torch.int32) - ref.to(torch.int32)).abs().max())
    print(f"u16 540p->360p: max code diff {diff}", flush=True)
    assert diff <= 1


def test_exact_and_float64_routes(lib):
    rng = np.random.default_rng(4)
    x = torch.from_numpy(rng.uniform(0, 1, (2, 540, 960)).astype(np.float32)).to(DEV)
    exact = dh.Executor(0, "exact")
    hp, vp = _plans(exact, 960, 540, 640, 360, kernel=dh.SPLINE36)
    got, ref = _both(lib, exact, hp, vp, x)
    assert torch.equal(got, ref)
    auto = dh.Executor(0, "auto")
    for sizes in ((960, 540, 700, 400), (640, 360, 448, 252)):
        h64, v64 = _plans(auto, *sizes, f64_mode=2)
        xs = x[:, :sizes[1], :sizes[0]].contiguous()
        got, ref = _both(lib, auto, h64, v64, xs)
        ulps = int(ulp_distance(got.cpu().numpy(), ref.cpu().numpy()).max())
        print(f"Float64 route {sizes}: max ULP vs allocating entry {ulps}", flush=True)
        assert ulps <= 1
        status, message = _call(lib, "dsmvc_hip_inverse_2d_batch_noalloc", auto, h64, v64, xs,
                                torch.empty_like(got), flags=_lib.ASYNC)
        assert status == _lib.INVALID_ARGUMENT and "synchronously" in message


def test_refusals(lib):
    ex = dh.Executor(0, "auto")
    hp, vp = _plans(ex, 1920, 1080, 1280, 720)
    big = torch.zeros((16, 1080, 1920), device=DEV)
    out = torch.empty((16, 720, 1280), device=DEV)
    status, message = _call(lib, "dsmvc_hip_inverse_2d_batch_noalloc", ex, hp, vp, big, out)
    assert status == _lib.INVALID_ARGUMENT and "scratch bytes" in message
    fresh_h = dh.Plan(dh.AxisRequest(1920, 1278, 1278.0))
    status, message = _call(lib, "dsmvc_hip_inverse_2d_batch_noalloc", ex, fresh_h, vp,
                            big[:1], torch.empty((1, 720, 1278), device=DEV))
    assert status == _lib.INVALID_ARGUMENT and "prepared" in message
    refined = dh.Executor(0, "refined")
    refined.prepare(hp)
    refined.prepare(vp)
    status, message = _call(lib, "dsmvc_hip_inverse_2d_batch_noalloc", refined, hp, vp, big[:1],
                            out[:1])
    assert status == _lib.INVALID_ARGUMENT and "refined" in message

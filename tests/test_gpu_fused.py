"""GPU parity of the fused 2D route (csrc/kernels_operator.hip): X = M_v Y M_h^T
with the explicit operators M = G^-1 A^T of the two plans multiplied on MFMA
in 16-bit split parts, no horizontal result in memory.  F32 planes run on
FP16 parts and fall back to an exact bfloat16 pass for samples beyond FP16's
range.

Bars, as for every Float32 route and integer plane of this library:
  * F32 planes: max-abs <= f32_bar(e_ref) against the FP64 oracle (the oracle's
    Float64 route on f64mode = 2 plans), 2e-6 at the workloads;
  * u8/u16 planes: <= 1 code against the oracle's scalar integer route.
Cases: the GRAYS F32 Debicubic 1080p -> 720p workload, all three planes of the
YUV420P16 workload in one planes call, odd geometries with several kernels,
pitched and unaligned rows (the kernel's element-wise staging), frame batches,
and plans whose operators do not fit (the call falls back to the two passes).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import (BICUBIC, BILINEAR, LANCZOS, SPLINE36, f32_bar, fp64_request,
                     make_request)  # noqa: E402
from workloads import integer_conversion, workload_plan_requests  # noqa: E402

DEV = torch.device("cuda:0")
POOL = 8


def _plan(r):
    return dh.Plan(dh.AxisRequest(r.source_size, r.destination_size, r.active_length, r.shift,
                                  r.kernel, r.taps, r.b, r.c, r.blur, r.border, r.f64_mode))


def _dconv(conv):
    return dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                conv.output_offset, conv.output_maximum)


@pytest.fixture(scope="module")
def fused():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return dh.Executor(0, "fused")


def _fp64(oracle, h, v, x):
    return oracle.inverse_2d_f64_exact(oracle.build(fp64_request(h)), oracle.build(fp64_request(v)),
                                       x.astype(np.float64))


def _check_f32(oracle, h, v, frames, got, label):
    def one(f):
        ref = _fp64(oracle, h, v, frames[f])
        ordered = oracle.inverse_2d(oracle.build(h), oracle.build(v), frames[f], ordered=True)
        return (float(np.abs(got[f].astype(np.float64) - ref).max()),
                float(np.abs(ordered.astype(np.float64) - ref).max()))
    with ThreadPoolExecutor(POOL) as pool:
        scores = list(pool.map(one, range(len(frames))))
    for f, (err, ref_err) in enumerate(scores):
        print(f"{label} frame {f}: max-abs vs FP64 {err:.3g} (reference F32 {ref_err:.3g})", flush=True)
        assert err <= f32_bar(ref_err), (label, f)


def test_operators_of_the_workload_fit():
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    for r in reqs:
        op = _plan(r).operator()
        assert op["rows_feasible"] and op["columns_feasible"], r
        assert op["steps"] <= 3 and op["dropped"] <= 2.0 ** -28


def test_grays_f32_debicubic_1080p_720p(fused, oracle):
    h, v = workload_plan_requests("yuv420p16_debicubic_1080p_720p")[:2]
    frames = np.random.default_rng(41).uniform(0, 1, (4, 1080, 1920)).astype(np.float32)
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(frames).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, frames, got, "grays f32 debicubic")


def test_yuv420p16_planes_call_within_one_code(fused, oracle):
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    rng = np.random.default_rng(42)
    jobs, hosts, convs, plans = [], [], [], []
    for p, (sw, sh) in enumerate([(1920, 1080), (960, 540), (960, 540)]):
        h, v = reqs[0:2] if p == 0 else reqs[2:4]
        x = rng.integers(0, 65536, (3, sh, sw)).astype(np.uint16)
        conv = integer_conversion(16, chroma=p > 0, full_range=False)
        xt = torch.from_numpy(x.astype(np.int32)).to(DEV).to(torch.uint16)
        jobs.append((_plan(h), _plan(v), xt, _dconv(conv)))
        hosts.append(x)
        convs.append(conv)
        plans.append((h, v))
    outs = fused.inverse_2d_planes(jobs)
    for p, out in enumerate(outs):
        got = out.to(torch.int32).cpu().numpy()
        h, v = plans[p]
        for f in range(3):
            want = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), hosts[p][f], convs[p])
            diff = int(np.abs(got[f] - want.astype(np.int32)).max())
            print(f"plane {p} frame {f}: max code diff {diff}", flush=True)
            assert diff <= 1


GEOMETRIES = [
    # (source w, source h, destination w, destination h, kernel, taps, shift)
    (333, 217, 250, 150, BILINEAR, 3, 0.0),
    (400, 300, 256, 171, LANCZOS, 3, 0.0),
    (640, 360, 427, 240, SPLINE36, 3, 0.25),
    (1280, 720, 853, 480, BICUBIC, 3, 0.0),
    (130, 70, 64, 33, BICUBIC, 3, 0.125),
    (97, 41, 17, 9, BILINEAR, 3, 0.0),
]


@pytest.mark.parametrize("geometry", GEOMETRIES, ids=lambda g: f"{g[0]}x{g[1]}-{g[2]}x{g[3]}-k{g[4]}")
def test_odd_geometries_f32_pitched(fused, oracle, geometry):
    sw, sh, dw, dh_, kernel, taps, shift = geometry
    h = make_request(sw, dw, float(dw), shift, kernel, taps)
    v = make_request(sh, dh_, float(dh_), 0.0, kernel, taps)
    rng = np.random.default_rng(sw * 7 + sh)
    frames = rng.uniform(0, 1, (3, sh, sw)).astype(np.float32)
    # Rows 3 samples longer than the plane: unaligned rows, element-wise staging.
    padded = np.zeros((3, sh, sw + 3), np.float32)
    padded[:, :, :sw] = frames
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(padded).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, frames, got, f"{sw}x{sh}->{dw}x{dh_}")


@pytest.mark.parametrize("bits", [8, 16])
def test_integer_odd_strides_within_one_code(fused, oracle, bits):
    h = make_request(721, 360, 360.0, 0.0, BICUBIC)
    v = make_request(405, 203, 203.0, 0.0, BICUBIC)
    dtype = np.uint8 if bits == 8 else np.uint16
    rng = np.random.default_rng(bits)
    x = rng.integers(0, 1 << bits, (2, 405, 721)).astype(dtype)
    conv = integer_conversion(bits, chroma=False, full_range=bits == 8)
    tdtype = torch.uint8 if bits == 8 else torch.uint16
    xt = torch.from_numpy(x.astype(np.int32)).to(DEV).to(tdtype)
    got = fused.inverse_2d(_plan(h), _plan(v), xt, conversion=_dconv(conv)).to(torch.int32).cpu().numpy()
    for f in range(2):
        want = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x[f], conv)
        diff = int(np.abs(got[f] - want.astype(np.int32)).max())
        print(f"u{bits} frame {f}: max code diff {diff}", flush=True)
        assert diff <= 1


def test_f32_values_beyond_fp16_take_the_exact_pass(fused, oracle):
    """F32 planes on the fused route are computed with FP16 parts first; a
    sample too large for that (here up to 3e5, against a bound of about
    3e4) sets the call's flag and the bfloat16 pass recomputes the call.
    Bar: f32_bar of the reference F32 path's own error on the same data."""
    h, v = make_request(640, 427, 427.0, 0.0, BICUBIC), make_request(360, 240, 240.0, 0.0, BICUBIC)
    rng = np.random.default_rng(45)
    frames = (rng.uniform(0, 1, (2, 360, 640)) * 3.0e5).astype(np.float32)
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(frames).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, frames, got, "f32 up to 3e5")
    # The same call on in-range data afterwards: the flag was cleared.
    small = (frames / 3.0e5).astype(np.float32)
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(small).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, small, got, "f32 in range after")


def test_plans_that_do_not_fit_fall_back(fused, oracle):
    """A 4:1 bicubic downscale: the operator's rows span about 130 source
    samples, more than the fused kernel stages; the call still solves within
    the bar (on the two-pass route)."""
    h = make_request(1920, 480, 480.0, 0.0, BICUBIC)
    v = make_request(270, 135, 135.0, 0.0, BICUBIC)
    assert not _plan(h).operator()["rows_feasible"]
    frames = np.random.default_rng(43).uniform(0, 1, (2, 270, 1920)).astype(np.float32)
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(frames).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, frames, got, "bicubic 4:1")


def test_near_identity_lanczos4_blur09(fused, oracle):
    """Delanczos taps=4 blur 0.9, 1920x1080 -> 1904x1071 on F32-only plans:
    ill-conditioned (rcond about 0.02), yet the FP64-built operator's band
    fits; bar f32_bar of the reference F32 path's own error."""
    h = make_request(1920, 1904, 1904.0, 0.0, LANCZOS, 4, blur=0.9, f64_mode=1)
    v = make_request(1080, 1071, 1071.0, 0.0, LANCZOS, 4, blur=0.9, f64_mode=1)
    assert _plan(h).operator()["rows_feasible"] and _plan(v).operator()["columns_feasible"]
    frames = np.random.default_rng(44).uniform(0, 1, (2, 1080, 1920)).astype(np.float32)
    got = fused.inverse_2d(_plan(h), _plan(v), torch.from_numpy(frames).to(DEV)).cpu().numpy()
    _check_f32(oracle, h, v, frames, got, "lanczos4 blur 0.9 near-identity")

"""GPU parity: the HIP executor (through the C ABI) against the oracle.

Bars (stated per test):
  * exact route (sequential solve): bit-identical to the ordered-FMA oracle
    (reference src/axis_plan.cpp:1068-1122 restated in oracle/);
  * parallel F32 routes (chunked, windowed, auto): max-abs against the FP64
    oracle (the oracle's Float64 route on the f64mode = 2 twin of each
    request) <= support.f32_bar(e_ref) = max(2e-6, 2 * e_ref), e_ref being the
    reference's ordered F32 path's own error on the same input;
  * integer output: <= 1 code vs the oracle's scalar integer route;
  * Float64 route: <= 1 output ULP vs the oracle's Float64 route;
  * padding outside the logical width is never written (guard sentinels).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import (BICUBIC, BILINEAR, F64_ALWAYS, F64_AUTO, F64_NEVER, LANCZOS, REFLECT101,
                     REPEAT, SPLINE16, SPLINE36, SPLINE64, SYMMETRIC, ZERO, f32_bar,
                     fp64_request, make_request, ulp_distance)  # noqa: E402
from workloads import workload_plan_requests, integer_conversion  # noqa: E402

DEV = torch.device("cuda:0")


def _hreq(r):
    return dh.AxisRequest(r.source_size, r.destination_size, r.active_length, r.shift,
                          r.kernel, r.taps, r.b, r.c, r.blur, r.border, r.f64_mode)


@pytest.fixture(scope="module")
def executors():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return {m: dh.Executor(0, m) for m in ("exact", "auto", "windowed")}


def _fp64_axis(oracle, req, x):
    """FP64 oracle of one axis over the vectors x [V][m]: [V][n] float64."""
    p = oracle.build(fp64_request(req))
    return np.stack([oracle.axis_f64(p, row.astype(np.float64)) for row in x])


def _check_fp64(got, ref64, ordered, label=""):
    """got and ordered (the reference's F32 result) against the FP64 oracle:
    max-abs <= f32_bar(the reference's own error)."""
    err = float(np.abs(got.astype(np.float64) - ref64).max())
    e_ref = float(np.abs(ordered.astype(np.float64) - ref64).max())
    assert err <= f32_bar(e_ref), (label, err, e_ref)
    return err


def _identical(a, b):
    """Bit-identical floats, except that +0 and -0 compare equal (the GPU
    applies zero-coefficient terms the reference's loops skip, which can only
    change the sign of an exact zero)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((a.view(np.uint32) == b.view(np.uint32))
                                              | ((a == 0) & (b == 0))))


AXIS_CASES = [
    ("bilinear", make_request(41, 35, 35.0, 0.1875, BILINEAR, border=SYMMETRIC, f64_mode=F64_NEVER)),
    ("bicubic", make_request(47, 39, 39.0, -0.3125, BICUBIC, border=REFLECT101, f64_mode=F64_NEVER)),
    ("lanczos3", make_request(59, 49, 49.0, 0.4375, LANCZOS, 3, border=REPEAT, f64_mode=F64_NEVER)),
    ("spline64", make_request(67, 57, 57.0, -0.25, SPLINE64, border=ZERO, f64_mode=F64_NEVER)),
    ("spline16-1080", make_request(1080, 720, kernel=SPLINE16, f64_mode=F64_NEVER)),
    ("spline36-1920", make_request(1920, 1280, kernel=SPLINE36, f64_mode=F64_NEVER)),
    ("bicubic-even-band", make_request(9, 3, kernel=BICUBIC, f64_mode=F64_NEVER)),
    ("lanczos-wide-blur", make_request(200, 150, kernel=LANCZOS, taps=3, blur=2.2, f64_mode=F64_NEVER)),
]


@pytest.mark.parametrize("name,req", AXIS_CASES, ids=[c[0] for c in AXIS_CASES])
def test_exact_rows_and_columns_bit_identical(executors, oracle, name, req):
    print(f"case {name}", flush=True)
    op = oracle.build(req)
    plan = dh.Plan(_hreq(req))
    m, n = req.source_size, req.destination_size
    rng = np.random.default_rng(m * 7 + n)
    vectors = 97
    # rows: input [vectors][m + 5] with guards, output [vectors][n + 3] pre-filled
    x = rng.uniform(-1, 1, (vectors, m + 5)).astype(np.float32)
    expected = np.stack([oracle.axis_f32_ordered(op, row[:m]) for row in x])
    out = torch.full((vectors, n + 3), -23.0, device=DEV)
    executors["exact"].inverse_rows(plan, torch.from_numpy(x).to(DEV), out)
    got = out.cpu().numpy()
    assert np.all(got[:, n:] == -23.0), "row padding was written"
    assert _identical(got[:, :n], expected)
    # columns: input [m][vectors + 3]
    xc = rng.uniform(-1, 1, (m, vectors + 3)).astype(np.float32)
    expected_c = np.stack([oracle.axis_f32_ordered(op, xc[:, v]) for v in range(vectors)], axis=1)
    outc = torch.full((n, vectors + 3), -31.0, device=DEV)
    executors["exact"].inverse_columns(plan, torch.from_numpy(xc).to(DEV)[:, :vectors], outc[:, :vectors])
    gotc = outc.cpu().numpy()
    assert np.all(gotc[:, vectors:] == -31.0), "column padding was written"
    assert _identical(gotc[:, :vectors], expected_c)


@pytest.mark.parametrize("name,req", AXIS_CASES, ids=[c[0] for c in AXIS_CASES])
def test_auto_rows_and_columns_close(executors, oracle, name, req):
    """Production (auto) route, rows and columns: max-abs against the FP64
    oracle <= f32_bar(e_ref)."""
    print(f"case {name}", flush=True)
    op = oracle.build(req)
    plan = dh.Plan(_hreq(req))
    m, n = req.source_size, req.destination_size
    rng = np.random.default_rng(m * 3 + n)
    vectors = 130
    x = rng.uniform(0, 1, (vectors, m)).astype(np.float32)
    expected = np.stack([oracle.axis_f32_ordered(op, row) for row in x])
    ref64 = _fp64_axis(oracle, req, x)
    got = executors["auto"].inverse_rows(plan, torch.from_numpy(x).to(DEV)).cpu().numpy()
    _check_fp64(got, ref64, expected, "rows")
    xc = np.ascontiguousarray(x.T)
    gotc = executors["auto"].inverse_columns(plan, torch.from_numpy(xc).to(DEV)).cpu().numpy()
    _check_fp64(gotc.T, ref64, expected, "columns")


@pytest.mark.parametrize("name,req", AXIS_CASES, ids=[c[0] for c in AXIS_CASES])
def test_windowed_rows_and_columns_close(executors, oracle, name, req):
    """Windowed (independent-chunk) route: max-abs against the FP64 oracle <=
    f32_bar(e_ref).  Plans the windowed kernels cannot take must be refused,
    not solved."""
    print(f"case {name}", flush=True)
    plan = dh.Plan(_hreq(req))
    info = plan.window_info()
    m, n = req.source_size, req.destination_size
    rng = np.random.default_rng(m * 5 + n)
    # 130 vectors: a partial last group of 64; 40: fewer vectors than one group.
    for vectors in (130, 40):
        x = rng.uniform(0, 1, (vectors, m)).astype(np.float32)
        xt = torch.from_numpy(x).to(DEV)
        if not info["feasible"]:
            with pytest.raises(ValueError, match="windowed"):
                executors["windowed"].inverse_columns(plan, xt.T.contiguous())
            return
        op = oracle.build(req)
        expected = np.stack([oracle.axis_f32_ordered(op, row) for row in x])
        ref64 = _fp64_axis(oracle, req, x)
        if info["rows_feasible"]:
            got = executors["windowed"].inverse_rows(plan, xt).cpu().numpy()
            _check_fp64(got, ref64, expected, f"rows {vectors}")
        gotc = executors["windowed"].inverse_columns(plan, xt.T.contiguous()).cpu().numpy()
        _check_fp64(gotc.T, ref64, expected, f"columns {vectors}")


LONG_CASES = [
    ("bicubic-3840-2000", make_request(3840, 2000, kernel=BICUBIC, f64_mode=F64_NEVER)),
    ("mitchell-1920-1778", make_request(1920, 1778, kernel=BICUBIC, b=1 / 3, c=1 / 3,
                                        f64_mode=F64_NEVER)),
    ("spline36-2160-1500", make_request(2160, 1500, kernel=SPLINE36, f64_mode=F64_NEVER)),
]


@pytest.mark.parametrize("name,req", LONG_CASES, ids=[c[0] for c in LONG_CASES])
def test_long_vectors_auto_close(executors, oracle, name, req):
    """Vectors longer than the carry-chain kernels take (n > 1280) run on the
    windowed kernels in auto mode: max-abs against the FP64 oracle <=
    f32_bar(e_ref), rows and columns, with enough vectors to fill the GPU."""
    print(f"case {name}", flush=True)
    plan = dh.Plan(_hreq(req))
    assert plan.window_info()["feasible"] == 1
    m, n = req.source_size, req.destination_size
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 1, (1024, m)).astype(np.float32)
    xt = torch.from_numpy(x).to(DEV)
    got = executors["auto"].inverse_rows(plan, xt).cpu().numpy()
    gotc = executors["auto"].inverse_columns(plan, xt.T.contiguous()).cpu().numpy()
    op = oracle.build(req)
    picks = [0, 1, 511, 1023]
    expected = np.stack([oracle.axis_f32_ordered(op, x[r]) for r in picks])
    ref64 = _fp64_axis(oracle, req, x[picks])
    _check_fp64(got[picks], ref64, expected, "rows")
    _check_fp64(gotc[:, picks].T, ref64, expected, "columns")


def test_windowed_yuv420p16_within_one_code(executors, oracle):
    """YUV420P16 Debicubic 1080p -> 720p geometry on the windowed kernels: <= 1 code vs the oracle's
    scalar integer route."""
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    rng = np.random.default_rng(22)
    for plane, (h, v, sw, sh) in enumerate([(reqs[0], reqs[1], 1920, 1080),
                                            (reqs[2], reqs[3], 960, 540)]):
        x = rng.integers(0, 65536, (2, sh, sw)).astype(np.uint16)
        conv = integer_conversion(16, chroma=plane > 0, full_range=False)
        dconv = dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                     conv.output_offset, conv.output_maximum)
        xt = torch.from_numpy(x.astype(np.int32)).to(DEV).to(torch.uint16)
        got = executors["windowed"].inverse_2d(dh.Plan(_hreq(h)), dh.Plan(_hreq(v)), xt,
                                               conversion=dconv).to(torch.int32).cpu().numpy()
        for f in range(2):
            expected = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x[f], conv)
            diff = int(np.abs(got[f] - expected.astype(np.int32)).max())
            print(f"plane {plane} frame {f}: max code diff {diff}", flush=True)
            assert diff <= 1


@pytest.mark.parametrize("kernel", [BILINEAR, BICUBIC, LANCZOS, SPLINE36, SPLINE64])
def test_exact_2d_bit_identical(executors, oracle, kernel):
    print(f"case kernel={kernel}", flush=True)
    h = make_request(480, 320, kernel=kernel, f64_mode=F64_NEVER)
    v = make_request(270, 180, kernel=kernel, f64_mode=F64_NEVER)
    x = np.random.default_rng(kernel).uniform(0, 1, (270, 480)).astype(np.float32)
    expected = oracle.inverse_2d(oracle.build(h), oracle.build(v), x, ordered=True)
    got = executors["exact"].inverse_2d(dh.Plan(_hreq(h)), dh.Plan(_hreq(v)),
                                        torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert _identical(got, expected)


def test_debilinear_1080p_720p_against_fp64(executors, oracle):
    """Debilinear geometry (Debilinear 1920x1080 -> 1280x720, 8 U[0,1] frames).
    Bars: exact == ordered oracle bit for bit; auto within 2e-6 of the FP64
    oracle (forced-F64 plans solved in double)."""
    reqs = workload_plan_requests("debilinear_1080p_720p")
    h, v = reqs
    rng = np.random.default_rng(1)
    frames = rng.uniform(0, 1, (8, 1080, 1920)).astype(np.float32)
    hp, vp = dh.Plan(_hreq(h)), dh.Plan(_hreq(v))
    xt = torch.from_numpy(frames).to(DEV)
    auto = executors["auto"].inverse_2d(hp, vp, xt).cpu().numpy()
    exact = executors["exact"].inverse_2d(hp, vp, xt).cpu().numpy()
    oh, ov = oracle.build(h), oracle.build(v)
    hq, vq = make_request(1920, 1280, kernel=BILINEAR, f64_mode=F64_ALWAYS), make_request(
        1080, 720, kernel=BILINEAR, f64_mode=F64_ALWAYS)
    fh, fv = oracle.build(hq), oracle.build(vq)
    for f in (0, 7):
        ordered = oracle.inverse_2d(oh, ov, frames[f], ordered=True)
        assert _identical(exact[f], ordered)
        ref64 = oracle.inverse_2d_f64_exact(fh, fv, frames[f].astype(np.float64))
        err = float(np.abs(auto[f].astype(np.float64) - ref64).max())
        print(f"frame {f}: auto vs FP64 max-abs {err:.3g}", flush=True)
        assert err <= 2e-6


def test_yuv420p16_debicubic_within_one_code(executors, oracle):
    """YUV420P16 geometry: Debicubic(0, 0.5) YUV420P16 1920x1080 -> 1280x720,
    limited range.  Bar: <= 1 code vs the oracle's scalar integer route."""
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    rng = np.random.default_rng(2)
    for plane, (h, v, sw, sh) in enumerate([(reqs[0], reqs[1], 1920, 1080),
                                            (reqs[2], reqs[3], 960, 540)]):
        x = rng.integers(0, 65536, (2, sh, sw)).astype(np.uint16)
        conv = integer_conversion(16, chroma=plane > 0, full_range=False)
        dconv = dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                     conv.output_offset, conv.output_maximum)
        xt = torch.from_numpy(x.astype(np.int32)).to(DEV).to(torch.uint16)
        for mode in ("exact", "auto"):
            got = executors[mode].inverse_2d(dh.Plan(_hreq(h)), dh.Plan(_hreq(v)), xt,
                                             conversion=dconv)
            got = got.to(torch.int32).cpu().numpy()
            for f in range(2):
                expected = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x[f], conv)
                diff = int(np.abs(got[f] - expected.astype(np.int32)).max())
                print(f"plane {plane} mode {mode} frame {f}: max code diff {diff}", flush=True)
                assert diff <= 1


def test_u8_integer_within_one_code(executors, oracle):
    h = make_request(300, 200, kernel=SPLINE36, f64_mode=F64_NEVER)
    v = make_request(170, 110, kernel=SPLINE36, f64_mode=F64_NEVER)
    x = np.random.default_rng(8).integers(0, 256, (170, 300)).astype(np.uint8)
    conv = integer_conversion(8, chroma=False, full_range=True)
    dconv = dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                 conv.output_offset, conv.output_maximum)
    expected = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x, conv).astype(np.int32)
    for mode in ("exact", "auto"):
        got = executors[mode].inverse_2d(dh.Plan(_hreq(h)), dh.Plan(_hreq(v)),
                                         torch.from_numpy(x).to(DEV), conversion=dconv)
        assert int(np.abs(got.cpu().numpy().astype(np.int32) - expected).max()) <= 1


def test_float64_route_within_one_ulp(executors, oracle):
    """Ill-conditioned Lanczos2 1080 -> 980 (auto-F64) mixed with a safe axis,
    and forced F64 on both axes.  Bar: <= 1 ULP vs the oracle's Float64 route."""
    cases = [
        (make_request(200, 150, kernel=BICUBIC, f64_mode=F64_NEVER),
         make_request(1080, 980, 978.1, 0.95, LANCZOS, 2, f64_mode=F64_AUTO)),
        (make_request(300, 250, kernel=LANCZOS, taps=4, blur=0.9, f64_mode=F64_ALWAYS),
         make_request(160, 100, kernel=SPLINE64, f64_mode=F64_ALWAYS)),
    ]
    for h, v in cases:
        x = np.random.default_rng(5).uniform(0, 1, (v.source_size, h.source_size)).astype(np.float32)
        expected = oracle.inverse_2d(oracle.build(h), oracle.build(v), x)
        got = executors["auto"].inverse_2d(dh.Plan(_hreq(h)), dh.Plan(_hreq(v)),
                                           torch.from_numpy(x).to(DEV)).cpu().numpy()
        assert int(ulp_distance(got, expected).max()) <= 1


def test_float64_windowed_route_matches_sequential(executors, oracle):
    """Float64 route on plans whose FP64 recurrence decays within the windowed
    FP64 kernels' warm-up (bicubic W=3, lanczos3 W=5, bilinear W=1 mixed with
    an F32 axis, spline64 and lanczos4 W=7 with 16-tap runs).  Bars: <= 1 ULP vs the oracle's Float64 route and vs the
    sequential Float64 kernel, with at most 1 in 1000 outputs differing from
    the sequential kernel at all."""
    cases = [
        (make_request(1920, 1280, kernel=BICUBIC, f64_mode=F64_ALWAYS),
         make_request(1080, 720, kernel=BICUBIC, f64_mode=F64_ALWAYS)),
        (make_request(1280, 1216, kernel=LANCZOS, taps=3, f64_mode=F64_ALWAYS),
         make_request(720, 684, kernel=LANCZOS, taps=3, f64_mode=F64_ALWAYS)),
        (make_request(640, 448, kernel=BILINEAR, f64_mode=F64_ALWAYS),
         make_request(360, 252, kernel=BICUBIC, f64_mode=F64_NEVER)),
        (make_request(960, 640, kernel=SPLINE64, f64_mode=F64_ALWAYS),
         make_request(540, 360, kernel=LANCZOS, taps=4, f64_mode=F64_ALWAYS)),
    ]
    for h, v in cases:
        hp, vp = dh.Plan(_hreq(h)), dh.Plan(_hreq(v))
        assert hp.window_info()["float64_feasible"] == 1
        x = np.random.default_rng(h.destination_size).uniform(
            0, 1, (2, v.source_size, h.source_size)).astype(np.float32)
        xt = torch.from_numpy(x).to(DEV)
        seq = executors["exact"].inverse_2d(hp, vp, xt).cpu().numpy()
        win = executors["auto"].inverse_2d(hp, vp, xt).cpu().numpy()
        flips = ulp_distance(win, seq)
        print(f"{h.destination_size}x{v.destination_size}: max ulp vs sequential "
              f"{int(flips.max())}, differing {float((flips > 0).mean()):.2e}", flush=True)
        assert int(flips.max()) <= 1 and float((flips > 0).mean()) <= 1e-3
        expected = oracle.inverse_2d(oracle.build(h), oracle.build(v), x[1])
        assert int(ulp_distance(win[1], expected).max()) <= 1


def test_float64_windowed_integer_within_one_code(executors, oracle):
    """Integer planes on the Float64 route (u16 and u8, forced FP64 plans the
    windowed FP64 kernels take): <= 1 code vs the oracle's integer Float64
    route (reference src/cpu_executor.cpp:1528-1651) and vs the sequential
    Float64 kernel."""
    cases = [(16, make_request(960, 640, kernel=BICUBIC, f64_mode=F64_ALWAYS),
              make_request(540, 360, kernel=BICUBIC, f64_mode=F64_ALWAYS)),
             (8, make_request(640, 480, kernel=LANCZOS, taps=3, f64_mode=F64_ALWAYS),
              make_request(360, 270, kernel=SPLINE36, f64_mode=F64_ALWAYS))]
    for bits, h, v in cases:
        hp, vp = dh.Plan(_hreq(h)), dh.Plan(_hreq(v))
        assert hp.window_info()["float64_feasible"] == 1 and vp.window_info()["float64_feasible"] == 1
        rng = np.random.default_rng(bits)
        dtype, tdtype = (np.uint16, torch.uint16) if bits == 16 else (np.uint8, torch.uint8)
        x = rng.integers(0, 1 << bits, (2, v.source_size, h.source_size)).astype(dtype)
        conv = integer_conversion(bits, chroma=False, full_range=bits == 8)
        dconv = dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                     conv.output_offset, conv.output_maximum)
        xt = torch.from_numpy(x.astype(np.int32)).to(DEV).to(tdtype)
        win = executors["auto"].inverse_2d(hp, vp, xt, conversion=dconv).to(torch.int32).cpu().numpy()
        seq = executors["exact"].inverse_2d(hp, vp, xt, conversion=dconv).to(torch.int32).cpu().numpy()
        print(f"u{bits}: max code diff vs sequential {int(np.abs(win - seq).max())}", flush=True)
        assert int(np.abs(win - seq).max()) <= 1
        for f in range(2):
            expected = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x[f], conv)
            assert int(np.abs(win[f] - expected.astype(np.int32)).max()) <= 1


def test_float64_windowed_route_is_taken(executors):
    """The automatic executor runs a 4K Float64 plane on the windowed FP64
    kernels: several times faster than the sequential Float64 kernel."""
    import time
    hp = dh.Plan(dh.AxisRequest(3840, 2048, 2048.0, f64_mode=2))
    vp = dh.Plan(dh.AxisRequest(2160, 1152, 1152.0, f64_mode=2))
    x = torch.rand(2160, 3840, device=DEV)
    elapsed = {}
    for mode in ("exact", "auto"):
        executors[mode].inverse_2d(hp, vp, x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            executors[mode].inverse_2d(hp, vp, x)
        torch.cuda.synchronize()
        elapsed[mode] = (time.perf_counter() - t0) / 3
    print({k: round(v * 1e3, 3) for k, v in elapsed.items()}, "ms", flush=True)
    assert elapsed["auto"] * 2 < elapsed["exact"]


def test_float64_route_rejects_nonfinite(executors):
    h = dh.Plan(dh.AxisRequest(64, 40, 40.0, f64_mode=2))
    v = dh.Plan(dh.AxisRequest(48, 30, 30.0, f64_mode=2))
    x = torch.rand(48, 64, device=DEV)
    x[10, 10] = float("nan")
    with pytest.raises(dh.DsmvcError, match="NaN or infinity"):
        executors["auto"].inverse_2d(h, v, x)


def test_batch_equals_single_frames(executors):
    h = dh.Plan(dh.AxisRequest(384, 256, 256.0))
    v = dh.Plan(dh.AxisRequest(216, 144, 144.0))
    x = torch.rand(5, 216, 384, device=DEV)
    for mode in ("exact", "auto"):
        batched = executors[mode].inverse_2d(h, v, x)
        for f in range(5):
            single = executors[mode].inverse_2d(h, v, x[f])
            assert torch.equal(batched[f], single)


def test_autotuned_partition_stable_and_close(oracle):
    """A batch that fills the GPU triggers the partition autotuner (chroma
    planes of the 1080p -> 720p Debicubic request, 64 frames).  Bars: the
    tuned executor repeats its result bit for bit; every partition the tuner
    can pick stays within f32_bar(e_ref) of the FP64 oracle."""
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    h, v = reqs[2], reqs[3]
    hp, vp = dh.Plan(_hreq(h)), dh.Plan(_hreq(v))
    x = torch.rand(64, 540, 960, device=DEV)
    tuned = dh.Executor(0, "auto")
    first = tuned.inverse_2d(hp, vp, x)
    second = tuned.inverse_2d(hp, vp, x)
    assert torch.equal(first, second)
    frames = x[[0, 63]].cpu().numpy()
    oh, ov = oracle.build(h), oracle.build(v)
    results = {"tuned": first[[0, 63]].cpu().numpy()}
    for count in (16, 8, 6):
        pinned = dh.Executor(0, "auto")
        pinned.tune(count, count)
        results[f"c{count}"] = pinned.inverse_2d(hp, vp, x[[0, 63]]).cpu().numpy()
    fh, fv = oracle.build(fp64_request(h)), oracle.build(fp64_request(v))
    for f in range(2):
        expected = oracle.inverse_2d(oh, ov, frames[f], ordered=True)
        ref64 = oracle.inverse_2d_f64_exact(fh, fv, frames[f].astype(np.float64))
        for name, got in results.items():
            err = _check_fp64(got[f], ref64, expected, name)
            print(f"{name} frame {f}: max-abs vs FP64 {err:.3g}", flush=True)


def test_planes_call_matches_per_plane_calls(executors, oracle):
    """dsmvc_hip_inverse_2d_planes: luma and chroma planes (u16, own plans
    and conversions) in one call equal the oracle within one code, and a
    float32 call mixing a windowed plane with a Float64-route plane equals
    the per-plane executor results (the Float64 plane bit for bit)."""
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    rng = np.random.default_rng(31)
    planes, jobs = [], []
    for index, (h, v, sw, sh) in enumerate([(reqs[0], reqs[1], 1920, 1080),
                                            (reqs[2], reqs[3], 960, 540),
                                            (reqs[2], reqs[3], 960, 540)]):
        x = rng.integers(0, 65536, (3, sh, sw)).astype(np.uint16)
        conv = integer_conversion(16, chroma=index > 0, full_range=False)
        dconv = dh.IntegerConversion(conv.input_offset, conv.input_scale, conv.output_scale,
                                     conv.output_offset, conv.output_maximum)
        planes.append((h, v, x, conv))
        jobs.append((dh.Plan(_hreq(h)), dh.Plan(_hreq(v)),
                     torch.from_numpy(x.astype(np.int32)).to(DEV).to(torch.uint16), dconv))
    outs = executors["auto"].inverse_2d_planes(jobs)
    for (h, v, x, conv), out in zip(planes, outs):
        got = out.to(torch.int32).cpu().numpy()
        for f in (0, 2):
            expected = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x[f], conv)
            assert int(np.abs(got[f] - expected.astype(np.int32)).max()) <= 1
    x = torch.rand(2, 270, 480, device=DEV)
    h64 = dh.Plan(_hreq(make_request(480, 320, kernel=BICUBIC, f64_mode=F64_ALWAYS)))
    v64 = dh.Plan(_hreq(make_request(270, 180, kernel=BICUBIC, f64_mode=F64_ALWAYS)))
    h32 = dh.Plan(_hreq(make_request(480, 300, kernel=LANCZOS, f64_mode=F64_NEVER)))
    v32 = dh.Plan(_hreq(make_request(270, 170, kernel=LANCZOS, f64_mode=F64_NEVER)))
    a, b = executors["auto"].inverse_2d_planes([(h64, v64, x, None), (h32, v32, x, None)])
    assert torch.equal(a, executors["auto"].inverse_2d(h64, v64, x))
    ref = executors["exact"].inverse_2d(h32, v32, x)
    assert float((b - ref).abs().max()) <= 4e-6 * max(1.0, float(ref.abs().max()))


def test_plugin_descale_yuv420(executors, oracle):
    """dsmvc_hip.Debicubic on a YUV420P16 clip equals the per-plane oracle
    within one code (chroma uses the shifted chroma plans)."""
    rng = np.random.default_rng(9)
    y = rng.integers(0, 65536, (1, 216, 384)).astype(np.uint16)
    u = rng.integers(0, 65536, (1, 108, 192)).astype(np.uint16)
    vv = rng.integers(0, 65536, (1, 108, 192)).astype(np.uint16)
    to = lambda a: torch.from_numpy(a.astype(np.int32)).to(DEV).to(torch.uint16)  # noqa: E731
    clip = dh.Clip([to(y), to(u), to(vv)], "yuv", 1, 1, 16, False)
    out = dh.Debicubic(clip, 256, 144, b=0.0, c=0.5)
    from workloads import chroma420_requests, luma_requests
    lreq = luma_requests(384, 216, 256, 144, BICUBIC)
    creq = chroma420_requests(384, 216, 256, 144, BICUBIC)
    for index, (plane, reqs) in enumerate([(y, lreq), (u, creq), (vv, creq)]):
        conv = integer_conversion(16, chroma=index > 0, full_range=False)
        expected = oracle.inverse_2d_int(oracle.build(reqs[0]), oracle.build(reqs[1]), plane[0],
                                         conv).astype(np.int32)
        got = out.planes[index][0].to(torch.int32).cpu().numpy()
        assert int(np.abs(got - expected).max()) <= 1

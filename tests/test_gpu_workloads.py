"""GPU parity at the benchmark workloads, scored against the FP64 oracle.

The FP64 oracle is the oracle's Float64 route on the same requests with
f64mode = 2 (plans that keep their FP64 factors; reference
src/axis_plan.cpp:909-940, 1193-1248).  Bars, written in each test:
  * Float32 routes: max-abs <= support.f32_bar(e_ref) = max(2e-6, 2 * e_ref),
    e_ref being the reference's ordered Float32 path's own error on the same
    input (2e-6 for every well-conditioned plan here);
  * Float64-route planes: <= 1 ULP of the oracle's Float64 route;
  * the 4:2:0 siting: ramps come back as ramps at the sited positions.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import (BICUBIC, BILINEAR, F64_ALWAYS, F64_AUTO, F64_NEVER, LANCZOS, MIRROR,
                     SPLINE36, SYMMETRIC, f32_bar, fp64_request, make_request,
                     ulp_distance)  # noqa: E402
from workloads import width_for_height, workload_plan_requests  # noqa: E402

DEV = torch.device("cuda:0")
POOL = 8   # oracle threads (ctypes releases the GIL during each call)


def _plan(r):
    return dh.Plan(dh.AxisRequest(r.source_size, r.destination_size, r.active_length, r.shift,
                                  r.kernel, r.taps, r.b, r.c, r.blur, r.border, r.f64_mode))


def _req(a: dh.AxisRequest):
    return make_request(a.source_size, a.destination_size, a.active_length, a.shift, a.kernel,
                        a.taps, a.b, a.c, a.blur, a.border, a.f64_mode)


@pytest.fixture(scope="module")
def executors():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return {m: dh.Executor(0, m) for m in ("exact", "auto", "windowed", "chunked")}


def _fp64_2d(oracle, h, v, x):
    return oracle.inverse_2d_f64_exact(oracle.build(fp64_request(h)),
                                       oracle.build(fp64_request(v)), x.astype(np.float64))


def _score(oracle, h, v, x, got):
    """(error vs FP64, reference F32 error vs FP64, error vs the ordered F32 oracle)."""
    ref64 = _fp64_2d(oracle, h, v, x)
    ordered = oracle.inverse_2d(oracle.build(h), oracle.build(v), x, ordered=True)
    return (float(np.abs(got.astype(np.float64) - ref64).max()),
            float(np.abs(ordered.astype(np.float64) - ref64).max()),
            float(np.abs(got - ordered).max()))


def test_grays_f32_debicubic_1080p_720p_against_fp64(executors, oracle):
    """The metric's workload: GRAYS F32 Debicubic(0, 0.5) 1920x1080 -> 1280x720,
    8 U[0,1] frames, on the production route (auto) and on each parallel
    family forced.  Bar: max-abs <= 2e-6 against the FP64 oracle."""
    h, v = workload_plan_requests("yuv420p16_debicubic_1080p_720p")[:2]
    hp, vp = _plan(h), _plan(v)
    frames = np.random.default_rng(1).uniform(0, 1, (8, 1080, 1920)).astype(np.float32)
    xt = torch.from_numpy(frames).to(DEV)
    results = {m: executors[m].inverse_2d(hp, vp, xt).cpu().numpy()
               for m in ("auto", "windowed", "chunked")}
    with ThreadPoolExecutor(POOL) as pool:
        refs = list(pool.map(lambda f: _fp64_2d(oracle, h, v, frames[f]), range(8)))
    for f in range(8):
        for mode, got in results.items():
            err = float(np.abs(got[f].astype(np.float64) - refs[f]).max())
            print(f"frame {f} {mode}: max-abs vs FP64 {err:.3g}", flush=True)
            assert err <= 2e-6, (f, mode)


def test_mitchell_501_plan_pairs_batch_against_oracle(executors, oracle):
    """The scan workload in full: Mitchell (b = c = 1/3) 1920x1080 -> h x
    round(16h/9) for every h in 500..1000, 501 plan pairs in one batch with
    graph replay.  Every pair is scored: max-abs <= f32_bar(e_ref) against
    the FP64 oracle, and its distance to the ordered F32 oracle is reported."""
    frame = np.random.default_rng(4).uniform(0, 1, (1080, 1920)).astype(np.float32)
    x = torch.from_numpy(frame).to(DEV)
    heights = list(range(500, 1001))
    pairs = []
    for hgt in heights:
        w = width_for_height(hgt)
        pairs.append((make_request(1920, w, kernel=BICUBIC, b=1 / 3, c=1 / 3),
                      make_request(1080, hgt, kernel=BICUBIC, b=1 / 3, c=1 / 3)))
    plans = dh.build_plans([dh.AxisRequest(r.source_size, r.destination_size, r.active_length,
                                           r.shift, r.kernel, r.taps, r.b, r.c)
                            for pair in pairs for r in pair])
    outs = [torch.full((vr.destination_size, hr.destination_size), float("nan"), device=DEV)
            for hr, vr in pairs]
    batch = dh.Batch(executors["auto"], [(plans[2 * k], plans[2 * k + 1], x, outs[k])
                                         for k in range(len(pairs))], graph=True)
    info = batch.info()
    print(info, flush=True)
    assert info["jobs"] == 501 and info["table_jobs"] == 501 and info["fallback_jobs"] == 0
    batch.run()
    got = [o.cpu().numpy() for o in outs]
    with ThreadPoolExecutor(POOL) as pool:
        scores = list(pool.map(lambda k: _score(oracle, *pairs[k], frame, got[k]),
                               range(len(pairs))))
    worst = max(range(len(pairs)), key=lambda k: scores[k][0])
    print(f"worst pair h={heights[worst]}: vs FP64 {scores[worst][0]:.3g}, reference F32 "
          f"{scores[worst][1]:.3g}, vs ordered {scores[worst][2]:.3g}", flush=True)
    print(f"max vs ordered F32 oracle over all pairs {max(s[2] for s in scores):.3g}", flush=True)
    for k, (err, e_ref, _) in enumerate(scores):
        assert np.isfinite(got[k]).all(), heights[k]
        assert err <= f32_bar(e_ref), (heights[k], err, e_ref)


KERNELS = [(BILINEAR, 3, 0.0, 0.5), (BICUBIC, 3, 0.0, 0.5), (LANCZOS, 3, 0.0, 0.5),
           (SPLINE36, 3, 0.0, 0.5)]


def test_served_stream_4k_1080p_every_plane_against_oracle(executors, oracle):
    """Served requests from 3840x2160, 1920x1080 and 1280x720 sources, gray
    and 4:2:0 YUV, every kernel and f64mode, submitted together.  Every plane
    is scored, chroma included (requests derived by filter_requests, the
    reference's prepare_filter_requests): Float64-route planes within 1 ULP of
    the oracle's Float64 route, Float32 planes within f32_bar(e_ref) of the
    FP64 oracle."""
    rng = np.random.default_rng(21)
    sources = [(3840, 2160), (1920, 1080), (1280, 720), (3840, 2160), (1920, 1080), (3840, 2160)]
    server = dh.Server(executors["auto"], memory_budget=1 << 30, slots=3)
    requests = []
    for k, (sw, sh) in enumerate(sources):
        hgt = int(sh * rng.uniform(0.5, 0.95)) & ~1
        w = width_for_height(hgt, even=True)
        kernel, taps, b, c = KERNELS[k % 4]
        mode = k % 3
        yuv = k % 2 == 0
        hreq, vreq = dh.filter_requests(sw, sh, w, hgt, 3 if yuv else 1, 1, 1, kernel=kernel,
                                        taps=taps, b=b, c=c, f64_mode=mode)
        planes = [(sw, sh, w, hgt, hreq[0], vreq[0])]
        if yuv:
            planes += [(sw // 2, sh // 2, w // 2, hgt // 2, hreq[1], vreq[1])] * 2
        ins = [rng.uniform(0, 1, (ph, pw)).astype(np.float32) for (pw, ph, *_r) in planes]
        tin = [torch.from_numpy(a).to(DEV) for a in ins]
        touts = [torch.full((dh_, dw), float("nan"), device=DEV) for (_, _, dw, dh_, *_r) in planes]
        rid = server.submit(tin, touts, kernel=kernel, taps=taps, b=b, c=c, f64_mode=mode)
        requests.append((rid, planes, ins, touts, mode))
    for rid, *_ in requests:
        server.wait(rid)
    jobs = []
    for rid, planes, ins, touts, mode in requests:
        for (pw, ph, dw, dh_, hr, vr), a, t in zip(planes, ins, touts):
            jobs.append((pw, ph, mode, _req(hr), _req(vr), a, t.cpu().numpy()))

    def score(job):
        pw, ph, mode, hr, vr, a, got = job
        oh, ov = oracle.build(hr), oracle.build(vr)
        if oh.c.requires_float64 or ov.c.requires_float64:
            expected = oracle.inverse_2d(oh, ov, a)
            return ("f64", int(ulp_distance(got, expected).max()), 0.0)
        err, e_ref, _ = _score(oracle, hr, vr, a, got)
        return ("f32", err, e_ref)

    with ThreadPoolExecutor(POOL) as pool:
        scores = list(pool.map(score, jobs))
    kinds = set()
    for (pw, ph, mode, hr, vr, *_), (route, err, e_ref) in zip(jobs, scores):
        kinds.add((pw >= 3840, route))
        print(f"{pw}x{ph} -> {hr.destination_size}x{vr.destination_size} f64mode {mode} "
              f"{route}: {err:.3g}", flush=True)
        if route == "f64":
            assert err <= 1
        else:
            assert err <= f32_bar(e_ref)
    # both routes ran at 4K
    assert (True, "f64") in kinds and (True, "f32") in kinds
    server.close()


def _forced(r, mode):
    return make_request(r.source_size, r.destination_size, r.active_length, r.shift, r.kernel,
                        r.taps, r.b, r.c, r.blur, r.border, mode)


FLOAT64_AXES = [
    ("forced-bicubic-1920", make_request(1920, 1280, kernel=BICUBIC, f64_mode=F64_ALWAYS)),
    ("auto-risk-lanczos2", make_request(1080, 980, 978.1, 0.95, LANCZOS, 2, f64_mode=F64_AUTO)),
    ("forced-lanczos4-blur", make_request(300, 250, kernel=LANCZOS, taps=4, blur=0.9,
                                          f64_mode=F64_ALWAYS)),
]


@pytest.mark.parametrize("name,req", FLOAT64_AXES, ids=[c[0] for c in FLOAT64_AXES])
def test_float64_single_axis_calls_within_one_ulp(executors, oracle, name, req):
    """Executor::inverse_rows / inverse_columns on plans that take the Float64
    route (reference src/cpu_executor.cpp:1417-1485: F32 samples in, FP64
    solve, F32 out).  Bar: <= 1 ULP of the oracle's Float64 route."""
    plan = _plan(req)
    assert plan.requires_float64
    op = oracle.build(req)
    m = req.source_size
    x = np.random.default_rng(m).uniform(0, 1, (70, m)).astype(np.float32)
    expected = oracle.rows(op, x)
    for mode in ("auto", "exact"):
        got = executors[mode].inverse_rows(plan, torch.from_numpy(x).to(DEV)).cpu().numpy()
        assert int(ulp_distance(got, expected).max()) <= 1, mode
        xc = np.ascontiguousarray(x.T)
        gotc = executors[mode].inverse_columns(plan, torch.from_numpy(xc).to(DEV)).cpu().numpy()
        assert int(ulp_distance(gotc, oracle.columns(op, xc)).max()) <= 1, mode


def test_mirror_padding_exact_and_auto(executors, oracle):
    """Padding mode 4 (the legacy mirror border, reference
    src/axis_plan.cpp:1060-1063): exact route bit-identical to the ordered
    oracle, auto route within f32_bar(e_ref) of the FP64 oracle."""
    cases = [(make_request(1920, 1280, kernel=BICUBIC, border=MIRROR, f64_mode=F64_NEVER),
              make_request(1080, 720, kernel=BICUBIC, border=MIRROR, f64_mode=F64_NEVER)),
             (make_request(640, 456, kernel=LANCZOS, taps=3, border=MIRROR, f64_mode=F64_NEVER),
              make_request(360, 250, kernel=SPLINE36, border=MIRROR, f64_mode=F64_NEVER))]
    for h, v in cases:
        x = np.random.default_rng(h.destination_size).uniform(
            0, 1, (v.source_size, h.source_size)).astype(np.float32)
        xt = torch.from_numpy(x).to(DEV)
        hp, vp = _plan(h), _plan(v)
        exact = executors["exact"].inverse_2d(hp, vp, xt).cpu().numpy()
        ordered = oracle.inverse_2d(oracle.build(h), oracle.build(v), x, ordered=True)
        same = (exact.view(np.uint32) == ordered.view(np.uint32)) | ((exact == 0) & (ordered == 0))
        assert bool(same.all())
        auto = executors["auto"].inverse_2d(hp, vp, xt).cpu().numpy()
        err, e_ref, _ = _score(oracle, h, v, x, auto)
        print(f"mirror {h.destination_size}x{v.destination_size}: auto vs FP64 {err:.3g} "
              f"(reference F32 {e_ref:.3g})", flush=True)
        assert err <= f32_bar(e_ref)


def test_chroma_siting_ramps_through_descale(executors):
    """Debilinear of a 4:2:0 float32 clip whose planes sample the ramp
    f(X, Y) = X / sw + Y / sh at their sited positions (luma at pixel
    centres, chroma left-sited horizontally and centre-sited vertically)
    returns that ramp at the destination's sited positions, on the GPU route,
    away from the borders.  Bar: 1e-5."""
    sw, sh, dw, dh_ = 1920, 1080, 1280, 720
    inner = 32

    def ramp(xs, ys):
        return (xs[None, :] / sw + ys[:, None] / sh).astype(np.float32)

    luma = ramp(np.arange(sw) + 0.5, np.arange(sh) + 0.5)
    chroma = ramp(2 * np.arange(sw // 2) + 0.5, 2 * np.arange(sh // 2) + 1.0)
    planes = [torch.from_numpy(p).to(DEV) for p in (luma, chroma, chroma)]
    out = dh.Debilinear(dh.Clip(planes, "yuv", 1, 1), dw, dh_)
    sx, sy = sw / dw, sh / dh_
    want_l = ramp((np.arange(dw) + 0.5) * sx, (np.arange(dh_) + 0.5) * sy)
    want_c = ramp((2 * np.arange(dw // 2) + 0.5) * sx, (2 * np.arange(dh_ // 2) + 1.0) * sy)
    for got, want in zip((p.cpu().numpy() for p in out.planes), (want_l, want_c, want_c)):
        err = float(np.abs(got - want)[inner:-inner, inner:-inner].max())
        print(f"ramp plane {got.shape}: interior max-abs {err:.3g}", flush=True)
        assert err <= 1e-5

"""GPU tests of multi-plan batches (include/dsmvc_hip.h, dsmvc_hip_batch_*).

Bars:
  * every table-solved job within f32_bar(e_ref) of the FP64 oracle's 2D
    solve, e_ref being the reference's ordered F32 path's own error (the same
    bar as any parallel F32 route);
  * a batch equals itself across runs bit for bit, with or without the
    captured HIP graph;
  * jobs the table kernels cannot take (Float64 route) are solved one by one
    and equal the executor's own result for that job bit for bit.
F32 jobs whose operators fit take the fused 2D route in the tables; the
two-pass executor (fused route off) keeps the windowed tables covered.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import BICUBIC, F64_ALWAYS, f32_bar, fp64_request, make_request  # noqa: E402
from workloads import width_for_height  # noqa: E402

DEV = torch.device("cuda:0")


def _plan(req):
    return dh.Plan(dh.AxisRequest(req.source_size, req.destination_size, req.active_length,
                                  req.shift, req.kernel, req.taps, req.b, req.c, req.blur,
                                  req.border, req.f64_mode))


@pytest.fixture(scope="module")
def executor():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return dh.Executor(0, "auto")


@pytest.fixture(scope="module")
def two_pass_executor():
    """Automatic mode with the fused 2D route off (DSMVC_HIP_FUSED=0 when the
    executor is made): batch jobs take the windowed tables or the fallback."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    old = os.environ.get("DSMVC_HIP_FUSED")
    os.environ["DSMVC_HIP_FUSED"] = "0"
    try:
        return dh.Executor(0, "auto")
    finally:
        if old is None:
            del os.environ["DSMVC_HIP_FUSED"]
        else:
            os.environ["DSMVC_HIP_FUSED"] = old


def _mitchell_jobs(heights, x):
    jobs = []
    for h in heights:
        w = width_for_height(h)
        hr = make_request(1920, w, kernel=BICUBIC, b=1 / 3, c=1 / 3)
        vr = make_request(1080, h, kernel=BICUBIC, b=1 / 3, c=1 / 3)
        out = torch.full((h, w), float("nan"), device=DEV)
        jobs.append((hr, vr, _plan(hr), _plan(vr), out))
    return jobs


@pytest.mark.parametrize("route", ["fused", "two-pass"])
def test_batch_of_uneven_plans_matches_oracle(route, executor, two_pass_executor, oracle):
    rng = np.random.default_rng(4)
    frame = rng.uniform(0, 1, (1080, 1920)).astype(np.float32)
    x = torch.from_numpy(frame).to(DEV)
    heights = [500, 537, 600, 719, 720, 811, 900, 999, 1000]
    jobs = _mitchell_jobs(heights, x)
    ex = executor if route == "fused" else two_pass_executor
    batch = dh.Batch(ex, [(hp, vp, x, out) for _, _, hp, vp, out in jobs])
    info = batch.info()
    print(route, info, flush=True)
    assert info["jobs"] == len(heights) and info["table_jobs"] == len(heights)
    # Fused jobs: two launches (the FP16 pass and the exact pass); windowed
    # jobs: one launch per (direction, band class, tap width, staging block
    # size) group, at most two block sizes per (direction, band class, tap
    # width).
    classes = {(rows, p.window_info()["band"], p.window_info()["width"])
               for _, _, hp, vp, _ in jobs for rows, p in ((True, hp), (False, vp))}
    assert info["fallback_jobs"] == 0 and info["launches"] <= 2 + 2 * len(classes)
    if route == "fused":
        assert info["fused_jobs"] >= 1
    else:
        assert info["fused_jobs"] == 0
    batch.run()
    for (hr, vr, _, _, out), h in zip(jobs, heights):
        got = out.cpu().numpy()
        assert np.isfinite(got).all()
        expected = oracle.inverse_2d(oracle.build(hr), oracle.build(vr), frame, ordered=True)
        ref64 = oracle.inverse_2d_f64_exact(oracle.build(fp64_request(hr)),
                                            oracle.build(fp64_request(vr)),
                                            frame.astype(np.float64))
        err = float(np.abs(got.astype(np.float64) - ref64).max())
        e_ref = float(np.abs(expected.astype(np.float64) - ref64).max())
        print(f"height {h}: max-abs vs FP64 {err:.3g} (reference F32 {e_ref:.3g})", flush=True)
        assert err <= f32_bar(e_ref)


def test_batch_runs_repeat_bit_for_bit_with_and_without_graph(executor):
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.uniform(0, 1, (1080, 1920)).astype(np.float32)).to(DEV)
    heights = list(range(500, 1001, 50))
    plain = _mitchell_jobs(heights, x)
    graphed = _mitchell_jobs(heights, x)
    a = dh.Batch(executor, [(hp, vp, x, out) for _, _, hp, vp, out in plain])
    b = dh.Batch(executor, [(hp, vp, x, out) for _, _, hp, vp, out in graphed], graph=True)
    assert b.info()["graph"] == 1
    a.run()
    first = [out.clone() for *_, out in plain]
    for replay in range(3):
        for *_, out in graphed:
            out.fill_(float("nan"))
        b.run()
        torch.cuda.synchronize()
        for (*_, out), ref in zip(graphed, first):
            assert torch.equal(out, ref), replay
        # Eager work and a synchronisation between replays.
        a.run()
        for (*_, out), ref in zip(plain, first):
            assert torch.equal(out, ref)


def test_batch_solves_float64_jobs_one_by_one(executor):
    rng = np.random.default_rng(6)
    x = torch.from_numpy(rng.uniform(0, 1, (2, 540, 960)).astype(np.float32)).to(DEV)
    hr = make_request(960, 640, kernel=BICUBIC, f64_mode=F64_ALWAYS)
    vr = make_request(540, 360, kernel=BICUBIC, f64_mode=F64_ALWAYS)
    hp, vp = _plan(hr), _plan(vr)
    fr = make_request(960, 700, kernel=BICUBIC)
    gr = make_request(540, 400, kernel=BICUBIC)
    fp, gp = _plan(fr), _plan(gr)
    out64 = torch.empty((2, 360, 640), device=DEV)
    out32 = torch.empty((2, 400, 700), device=DEV)
    batch = dh.Batch(executor, [(hp, vp, x, out64), (fp, gp, x, out32)])
    info = batch.info()
    assert info["fallback_jobs"] == 1 and info["table_jobs"] == 1
    batch.run()
    expected = executor.inverse_2d(hp, vp, x)
    assert torch.equal(out64, expected)


def test_batch_job_with_one_windowed_axis_runs_once_on_the_fallback(two_pass_executor, oracle):
    """A job whose horizontal plan the windowed kernels take but whose
    vertical plan they do not (band 9) is counted as a fallback job, solved
    once by the executor, and equals the executor's own result; a windowed
    job in the same batch still runs in the tables."""
    from support import LANCZOS, F64_NEVER
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.uniform(0, 1, (540, 960)).astype(np.float32)).to(DEV)
    hr = make_request(960, 640, kernel=BICUBIC)
    vr = make_request(540, 360, kernel=LANCZOS, taps=5, f64_mode=F64_NEVER)
    hp, vp = _plan(hr), _plan(vr)
    assert hp.window_info()["rows_feasible"] == 1 and vp.window_info()["feasible"] == 0
    gp = _plan(make_request(540, 400, kernel=BICUBIC))
    fp = _plan(make_request(960, 700, kernel=BICUBIC))
    mixed = torch.full((360, 640), float("nan"), device=DEV)
    plain = torch.full((400, 700), float("nan"), device=DEV)
    executor = two_pass_executor
    batch = dh.Batch(executor, [(hp, vp, x, mixed), (fp, gp, x, plain)])
    info = batch.info()
    assert info["fallback_jobs"] == 1 and info["table_jobs"] == 1
    batch.run()
    assert torch.equal(mixed, executor.inverse_2d(hp, vp, x))
    assert float((plain - executor.inverse_2d(fp, gp, x)).abs().max()) <= 2e-6

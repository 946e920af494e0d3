"""Precision on ill-conditioned plans: FP32, FP32 with refinement, and the
Float64 route, each scored against the FP64 oracle.

Workload: Delanczos taps=4 and Despline64 with blur 0.9 on GRAYS F32
1920x1080 -> 1904x1071 (near-identity, condition ~50) and -> 960x540, and
Delanczos taps=5 (band 9) -> 1904x1071, and Delanczos taps=4 -> 240x135; frames
of seeded uniform [0, 1] noise plus a sum of 8 random 2-D sinusoids.  The
oracle is the C restatement's Float64 route on forced-FP64 plans (reference
src/axis_plan.cpp:1124-1248 with f64mode=2).

Bars (max-abs over the frame):
  * Float64 route: <= 2 F32 ULPs of the output scale (it rounds once to F32);
  * FP32 + one refinement step: <= 2e-6, and no worse than plain FP32;
  * FP32 + two steps: no worse than one step (within one F32 ULP of scale).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dsmvc_hip as dh  # noqa: E402
from support import F64_ALWAYS, F64_NEVER, LANCZOS, SPLINE64, make_request  # noqa: E402

DEV = torch.device("cuda:0")

CASES = [
    ("lanczos4-1904x1071", LANCZOS, 4, 1904, 1071),
    ("spline64-1904x1071", SPLINE64, 3, 1904, 1071),
    ("lanczos4-960x540", LANCZOS, 4, 960, 540),
    ("spline64-960x540", SPLINE64, 3, 960, 540),
    # Band 9: beyond the windowed kernels (bands up to 7), so refinement runs
    # the one-lane-per-vector residual and correction kernels.
    ("lanczos5-1904x1071", LANCZOS, 5, 1904, 1071),
    # 8x downscale: about 58 taps per element, so the vertical residual's
    # four-element blocks sum element by element (their sample rows exceed the
    # LDS tile).
    ("lanczos4-240x135", LANCZOS, 4, 240, 135),
]


def frame_with_sinusoids(seed, w=1920, h=1080):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, (h, w))
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(8):
        amp = rng.uniform(0.05, 0.25)
        fx, fy = rng.uniform(0.002, 0.2, 2)
        phase = rng.uniform(0, 2 * np.pi)
        x += amp * np.sin(2 * np.pi * (fx * xx + fy * yy) + phase)
    return x.astype(np.float32)


def _plan(w_src, w_dst, kernel, taps, mode):
    return dh.Plan(dh.AxisRequest(w_src, w_dst, float(w_dst), 0.0, kernel, taps, 0.0, 0.5, 0.9,
                                  dh.SYMMETRIC, mode))


@pytest.fixture(scope="module")
def executors():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    refined1 = dh.Executor(0, "refined")
    refined2 = dh.Executor(0, "refined")
    refined2.set_refinement(2)
    return {"auto": dh.Executor(0, "auto"), "refined1": refined1, "refined2": refined2}


@pytest.mark.parametrize("name,kernel,taps,dw,dh_", CASES, ids=[c[0] for c in CASES])
def test_refinement_against_fp64_oracle(executors, oracle, name, kernel, taps, dw, dh_):
    frame = frame_with_sinusoids(len(name))
    x = torch.from_numpy(frame).to(DEV)
    h32, v32 = _plan(1920, dw, kernel, taps, F64_NEVER), _plan(1080, dh_, kernel, taps, F64_NEVER)
    h64, v64 = _plan(1920, dw, kernel, taps, F64_ALWAYS), _plan(1080, dh_, kernel, taps, F64_ALWAYS)
    assert h64.requires_float64 and v64.requires_float64
    if taps == 5:
        assert h64.half_bandwidth > 7 and v64.half_bandwidth > 7
    plain = executors["auto"].inverse_2d(h32, v32, x).cpu().numpy()
    f64 = executors["auto"].inverse_2d(h64, v64, x).cpu().numpy()
    ref1 = executors["refined1"].inverse_2d(h64, v64, x).cpu().numpy()
    ref2 = executors["refined2"].inverse_2d(h64, v64, x).cpu().numpy()
    oh = oracle.build(make_request(1920, dw, kernel=kernel, taps=taps, blur=0.9, f64_mode=F64_ALWAYS))
    ov = oracle.build(make_request(1080, dh_, kernel=kernel, taps=taps, blur=0.9, f64_mode=F64_ALWAYS))
    exact = oracle.inverse_2d_f64_exact(oh, ov, frame.astype(np.float64))
    scale = float(np.abs(exact).max())
    ulp = float(np.spacing(np.float32(scale)))
    err = {k: float(np.abs(v.astype(np.float64) - exact).max())
           for k, v in (("fp32", plain), ("fp32+refine1", ref1), ("fp32+refine2", ref2), ("f64", f64))}
    print(f"{name}: scale {scale:.3f} rcond h {h64.normal_rcond:.3g} v {v64.normal_rcond:.3g} "
          + " ".join(f"{k} {v:.3g}" for k, v in err.items()), flush=True)
    assert err["f64"] <= 2 * ulp
    assert err["fp32+refine1"] <= 2e-6
    assert err["fp32+refine1"] <= err["fp32"] + ulp
    assert err["fp32+refine2"] <= err["fp32+refine1"] + ulp


def test_refined_frame_batch_matches_single_frames(executors, oracle):
    """Refined 2D solves over a batch of pitched frames equal the same solves
    frame by frame, bit for bit, and stay within 2e-6 of the FP64 oracle.
    Delanczos taps=4 blur 0.9, 640x360 -> 600x338 (windowed, band 7)."""
    hp, vp = _plan(640, 600, LANCZOS, 4, F64_ALWAYS), _plan(360, 338, LANCZOS, 4, F64_ALWAYS)
    assert hp.window_info()["feasible"] and vp.window_info()["feasible"]
    frames = [frame_with_sinusoids(11 + f, 640, 360) for f in range(5)]
    x = torch.zeros((5, 360, 672), dtype=torch.float32, device=DEV)
    x[:, :, :640] = torch.from_numpy(np.stack(frames)).to(DEV)
    refined = executors["refined1"]
    batch = refined.inverse_2d(hp, vp, x)
    oh = oracle.build(make_request(640, 600, kernel=LANCZOS, taps=4, blur=0.9, f64_mode=F64_ALWAYS))
    ov = oracle.build(make_request(360, 338, kernel=LANCZOS, taps=4, blur=0.9, f64_mode=F64_ALWAYS))
    for f in range(5):
        single = refined.inverse_2d(hp, vp, x[f])
        assert torch.equal(single, batch[f]), f"frame {f}"
    for f in (0, 4):
        exact = oracle.inverse_2d_f64_exact(oh, ov, frames[f].astype(np.float64))
        err = float(np.abs(batch[f].cpu().numpy().astype(np.float64) - exact).max())
        print(f"frame {f}: refined max-abs {err:.3g}", flush=True)
        assert err <= 2e-6


@pytest.mark.parametrize("src,dst,vectors", [(640, 600, 300), (360, 338, 200), (20, 18, 3),
                                             (24, 22, 5)])
def test_refined_single_axis_against_fp64(executors, oracle, src, dst, vectors):
    """Refined rows and columns calls (tiny planes and vector counts included:
    fewer elements than one residual block, fewer vectors than a thread's
    share) against the FP64 oracle's Float64 axis solve, <= 2e-6."""
    plan = _plan(src, dst, LANCZOS, 4, F64_ALWAYS)
    assert plan.window_info()["feasible"]
    op = oracle.build(make_request(src, dst, kernel=LANCZOS, taps=4, blur=0.9, f64_mode=F64_ALWAYS))
    rng = np.random.default_rng(src + vectors)
    y = rng.uniform(0.0, 1.0, (vectors, src)).astype(np.float32)
    refined = executors["refined1"]
    rows = refined.inverse_rows(plan, torch.from_numpy(y).to(DEV)).cpu().numpy()
    want = np.stack([oracle.axis_f64(op, v) for v in y])
    assert float(np.abs(rows.astype(np.float64) - want).max()) <= 2e-6
    yc = np.ascontiguousarray(y.T)
    cols = refined.inverse_columns(plan, torch.from_numpy(yc).to(DEV)).cpu().numpy()
    assert float(np.abs(cols.astype(np.float64) - want.T).max()) <= 2e-6

"""CPU tests of the windowed (independent-chunk) partition.

The windowed kernels solve every chunk of L destination rows on its own: the
forward recurrence starts P rows early from a zero state, and the backward
recurrence starts P rows late from a zero state.  The library picks P per band
class and accepts a plan only when the influence of the wrong starting state
has decayed below 2^-28 of the carried values (`residual`).

These tests restate the windowed algorithm in float64 numpy, on the plan's own
F32 coefficients, and compare it with the full sequential solve of the same
coefficients.  The difference is then pure truncation, and must stay within
the residual the library reports.  A warm-up shorter than the library's must
fail the same bound, so the check is sensitive to a wrong boundary.
"""
import numpy as np
import pytest

import dsmvc_hip as dh
from support import BICUBIC, BILINEAR, LANCZOS, SPLINE36, SPLINE64, make_request
from workloads import workload_plan_requests

TOLERANCE = 2.0 ** -28


def _plan(req):
    return dh.Plan(dh.AxisRequest(req.source_size, req.destination_size, req.active_length,
                                  req.shift, req.kernel, req.taps, req.b, req.c, req.blur,
                                  req.border, req.f64_mode))


def _coefficients(plan):
    a = plan.arrays()
    n, w = plan.destination_size, plan.half_bandwidth
    lld = a["lower_ld"].astype(np.float64).reshape(w, n) if w else np.zeros((0, n))
    ul = a["upper_l"].astype(np.float64).reshape(w, n) if w else np.zeros((0, n))
    return a, lld, ul, a["inverse_diagonal"].astype(np.float64)


def _rhs(plan, y):
    """A^T y in float64 from the plan's CSR (reference axis_plan.cpp:972-1018)."""
    a = plan.arrays()
    off, idx, wt = a["offsets"], a["indices"], a["weights"].astype(np.float64)
    rows = np.repeat(np.arange(plan.destination_size), np.diff(off).astype(np.int64))
    out = np.zeros((y.shape[0], plan.destination_size))
    np.add.at(out.T, rows, (wt[:, None] * y[:, idx].T))
    return out


def _forward(b, lld, inv, start, stop, w):
    """Forward recurrence over rows [start, stop) from a zero state."""
    y = np.zeros((b.shape[0], stop - start))
    for i in range(start, stop):
        v = b[:, i].copy()
        for d in range(1, w + 1):
            if i - d >= start:
                v -= lld[d - 1, i] * y[:, i - d - start]
        y[:, i - start] = v * inv[i]
    return y


def _backward(y, ul, start, stop, w, offset):
    """Backward recurrence over rows [start, stop) from a zero state; y holds
    rows [offset, ...)."""
    x = np.zeros((y.shape[0], stop - start))
    for i in range(stop - 1, start - 1, -1):
        v = y[:, i - offset].copy()
        for d in range(1, w + 1):
            if i + d < stop:
                v -= ul[d - 1, i] * x[:, i + d - start]
        x[:, i - start] = v
    return x


def _sequential(b, lld, ul, inv, w):
    n = b.shape[1]
    return _backward(_forward(b, lld, inv, 0, n, w), ul, 0, n, w, 0)


def _windowed(b, lld, ul, inv, w, length, warmup):
    n = b.shape[1]
    out = np.zeros_like(b)
    for a in range(0, n, length):
        e = min(n, a + length)
        f0, f1 = max(0, a - warmup), min(n, e + warmup)
        y = _forward(b, lld, inv, f0, f1, w)
        x = _backward(y, ul, a, f1, w, f0)
        out[:, a:e] = x[:, : e - a]
    return out


CASES = [
    ("bilinear-1920-1280", make_request(1920, 1280, kernel=BILINEAR)),
    ("bicubic-1080-720", make_request(1080, 720, kernel=BICUBIC)),
    ("lanczos3-1920-1280", make_request(1920, 1280, kernel=LANCZOS, taps=3)),
    ("spline36-1280-1152", make_request(1280, 1152, kernel=SPLINE36)),
    ("spline64-1920-1280", make_request(1920, 1280, kernel=SPLINE64)),
    ("mitchell-1920-1778", make_request(1920, 1778, kernel=BICUBIC, b=1 / 3, c=1 / 3)),
    ("bicubic-3840-2000", make_request(3840, 2000, kernel=BICUBIC)),
]


@pytest.mark.parametrize("name,req", CASES, ids=[c[0] for c in CASES])
def test_windowed_solve_matches_sequential(name, req):
    plan = _plan(req)
    info = plan.window_info()
    assert info["feasible"] == 1, info
    assert info["residual"] <= TOLERANCE
    assert info["band"] in (1, 3, 5, 7) and info["band"] >= plan.half_bandwidth
    assert info["chunks"] == -(-plan.destination_size // info["length"])
    _, lld, ul, inv = _coefficients(plan)
    rng = np.random.default_rng(plan.destination_size)
    y = rng.uniform(0.0, 1.0, (4, plan.source_size))
    b = _rhs(plan, y)
    w = plan.half_bandwidth
    exact = _sequential(b, lld, ul, inv, w)
    scale = float(np.abs(exact).max())
    err = float(np.abs(_windowed(b, lld, ul, inv, w, info["length"], info["warmup"]) - exact).max())
    # The truncation is below the residual times the carried values (which
    # stay within a small multiple of the solution's scale).
    assert err <= 8 * TOLERANCE * scale, (err, scale)
    # A quarter of the warm-up leaves the wrong start visible.
    short = _windowed(b, lld, ul, inv, w, info["length"], info["warmup"] // 4)
    assert float(np.abs(short - exact).max()) > 8 * TOLERANCE * scale


def _streamed(b, lld, ul, inv, w, length, warmup, span):
    """The kernels' segments: each wavefront solves `span` consecutive chunks.
    The forward recurrence starts from a zero state at least 2P rows before
    the segment's first chunk ends its warm-up (whole chunks of L rows, the
    smallest count covering 2P: padded row (q0 - ceil(2P / L)) L + 2P, plan
    row minus P, clamped to 0) and runs on without a restart; each chunk's
    backward recurrence starts P rows past the chunk from a zero state."""
    n = b.shape[1]
    chunks = -(-n // length)
    warm = -(-2 * warmup // length)
    out = np.zeros_like(b)
    for q0 in range(0, chunks, span):
        q1 = min(q0 + span, chunks)
        f0 = max(0, (q0 - warm) * length + warmup)
        f1 = min(n, q1 * length + warmup)
        y = _forward(b, lld, inv, f0, f1, w)
        for q in range(q0, q1):
            a, e = q * length, min(n, (q + 1) * length)
            x = _backward(y, ul, a, min(n, e + warmup), w, f0)
            out[:, a:e] = x[:, : e - a]
    return out


@pytest.mark.parametrize("name,req", CASES[:4], ids=[c[0] for c in CASES[:4]])
@pytest.mark.parametrize("span", [1, 2, 3, 5])
def test_streamed_segments_match_sequential(name, req, span):
    """Segments of several chunks, as the kernels run them: the forward sweep
    carries from chunk to chunk, so only the segment's start and each chunk's
    backward end are truncated, and the same bound holds."""
    plan = _plan(req)
    info = plan.window_info()
    _, lld, ul, inv = _coefficients(plan)
    rng = np.random.default_rng(plan.destination_size + span)
    b = _rhs(plan, rng.uniform(0.0, 1.0, (2, plan.source_size)))
    w = plan.half_bandwidth
    exact = _sequential(b, lld, ul, inv, w)
    scale = float(np.abs(exact).max())
    got = _streamed(b, lld, ul, inv, w, info["length"], info["warmup"], span)
    assert float(np.abs(got - exact).max()) <= 8 * TOLERANCE * scale
    short = _streamed(b, lld, ul, inv, w, info["length"], info["warmup"] // 4, span)
    assert float(np.abs(short - exact).max()) > 8 * TOLERANCE * scale


@pytest.mark.parametrize("workload", ["debilinear_1080p_720p", "yuv420p16_debicubic_1080p_720p",
                                      "mitchell_21_heights", "served_mix_sample"])
def test_workload_plans_are_windowed(workload):
    """Every plan of the throughput workloads (W <= 7, taps <= 16) is
    accepted, in both directions."""
    for req in workload_plan_requests(workload):
        plan = _plan(req)
        info = plan.window_info()
        assert info["feasible"] == 1 and info["rows_feasible"] == 1, (req.source_size,
                                                                      req.destination_size, info)
        assert info["residual"] <= TOLERANCE


def test_near_identity_plans_report_their_decay():
    """The near-identity plans (1920 -> 1904, blur 0.9, W = 7) decay
    slowly: whatever the verdict, a feasible plan must meet the tolerance and
    an infeasible one must report a residual above it."""
    for req in workload_plan_requests("near_identity_blur09"):
        info = _plan(req).window_info()
        if info["feasible"]:
            assert info["residual"] <= TOLERANCE
        else:
            assert info["residual"] > TOLERANCE or info["width"] == 0


# ---- Float64 route --------------------------------------------------------
# The windowed FP64 kernels run the Float64 route's recurrence (unit-L FP64
# factors: z_i = b_i - sum_d L(i, i-d) z_{i-d}, y_i = z_i / (D_i + eps),
# x_i = y_i - sum_d L(i+d, i) x_{i+d}; reference src/axis_plan.cpp:1124-1191)
# on chunks of L rows with P warm-up rows per side, and accept a plan when
# the wrong start has decayed below 2^-40.  Chunk shape per band class:
F64_SHAPES = {1: (32, 32), 3: (32, 32), 5: (24, 48), 7: (24, 64)}
TOLERANCE64 = 2.0 ** -40


def _f64_factors(plan):
    a = plan.arrays()
    n, w = plan.destination_size, plan.half_bandwidth
    bands = a["ldlt_bands_f64"].reshape(w + 1, n)
    low, up = np.zeros((w, n)), np.zeros((w, n))
    for d in range(1, w + 1):
        low[d - 1, d:] = bands[d, : n - d]      # L(i, i - d), stored at column i - d
        up[d - 1, : n - d] = bands[d, : n - d]  # L(i + d, i), stored at column i
    return a, low, up, a["inverse_diagonal_f64"].astype(np.float64)


def _forward_z(b, low, inv, start, stop, w):
    """Unit-L forward recurrence over [start, stop) from a zero state; returns
    the diagonal-scaled values."""
    z = np.zeros((b.shape[0], stop - start))
    for i in range(start, stop):
        v = b[:, i].copy()
        for d in range(1, w + 1):
            if i - d >= start:
                v -= low[d - 1, i] * z[:, i - d - start]
        z[:, i - start] = v
    return z * inv[start:stop]


def _windowed64(b, low, up, inv, w, length, warmup):
    n = b.shape[1]
    out = np.zeros_like(b)
    for a in range(0, n, length):
        e = min(n, a + length)
        f0, f1 = max(0, a - warmup), min(n, e + warmup)
        y = _forward_z(b, low, inv, f0, f1, w)
        x = _backward(y, up, a, f1, w, f0)
        out[:, a:e] = x[:, : e - a]
    return out


F64_CASES = [
    ("bilinear-1920-1280", make_request(1920, 1280, kernel=BILINEAR, f64_mode=2)),
    ("bicubic-1080-756", make_request(1080, 756, kernel=BICUBIC, f64_mode=2)),
    ("lanczos3-2160-1512", make_request(2160, 1512, kernel=LANCZOS, taps=3, f64_mode=2)),
    ("spline36-1080-1026", make_request(1080, 1026, kernel=SPLINE36, f64_mode=2)),
    ("spline64-1920-1280", make_request(1920, 1280, kernel=SPLINE64, f64_mode=2)),
]


@pytest.mark.parametrize("name,req", F64_CASES, ids=[c[0] for c in F64_CASES])
def test_windowed_float64_solve_matches_sequential(name, req):
    plan = _plan(req)
    assert plan.requires_float64
    info = plan.window_info()
    assert info["float64_feasible"] == 1, info
    a, low, up, inv = _f64_factors(plan)
    w = plan.half_bandwidth
    length, warmup = F64_SHAPES[info["band"]]
    rng = np.random.default_rng(plan.source_size)
    y = rng.uniform(0.0, 1.0, (3, plan.source_size))
    off, idx = a["offsets"], a["indices"]
    rows = np.repeat(np.arange(plan.destination_size), np.diff(off).astype(np.int64))
    b = np.zeros((3, plan.destination_size))
    np.add.at(b.T, rows, a["weights_f64"].astype(np.float64)[:, None] * y[:, idx].T)
    exact = _backward(_forward_z(b, low, inv, 0, plan.destination_size, w), up, 0,
                      plan.destination_size, w, 0)
    scale = float(np.abs(exact).max())
    err = float(np.abs(_windowed64(b, low, up, inv, w, length, warmup) - exact).max())
    assert err <= 8 * TOLERANCE64 * scale, (err, scale)
    short = _windowed64(b, low, up, inv, w, length, warmup // 4)
    assert float(np.abs(short - exact).max()) > 8 * TOLERANCE64 * scale

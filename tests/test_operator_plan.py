"""The explicit axis operator of the fused 2D route (dsmvc_hip_plan_operator,
csrc/operator_plan.cpp), pinned on the CPU by what the mathematics fixes:

  * M = G^-1 A^T with G = A^T A, so M A = I (the descale of an exact
    upscale returns its input) and G M = A^T, up to the truncated tails;
  * A maps constants to constants (each source sample's taps sum to 1), so
    M 1 = 1;
  * M y equals the oracle's FP64 axis solve on the f64mode = 2 plan of the
    same request (reference src/axis_plan.cpp:1193-1248);
  * truncation drops at most 2^-28 of each row's L1 mass at either end, and
    the kept band is the contiguous run around the destination's position.
"""
import numpy as np
import pytest

import dsmvc_hip as dh
from support import BICUBIC, BILINEAR, LANCZOS, SPLINE36, fp64_request, make_request

CASES = [
    (1920, 1280, BICUBIC, 3, 0.0, 1.0),
    (1080, 720, BICUBIC, 3, 0.0, 1.0),
    (960, 640, BICUBIC, 3, 0.25 - 0.25 * 1280 / 1920, 1.0),
    (1920, 1280, BILINEAR, 3, 0.0, 1.0),
    (1280, 853, LANCZOS, 3, 0.0, 1.0),
    (640, 427, SPLINE36, 3, 0.125, 1.0),
    (333, 250, BICUBIC, 3, 0.0, 0.9),
]


def _plan(r):
    return dh.Plan(dh.AxisRequest(r.source_size, r.destination_size, r.active_length, r.shift,
                                  r.kernel, r.taps, r.b, r.c, r.blur, r.border, r.f64_mode))


def _dense(plan):
    a = plan.arrays()
    n, m = plan.destination_size, plan.source_size
    A = np.zeros((m, n))
    for i in range(n):
        s, e = a["offsets"][i], a["offsets"][i + 1]
        A[a["indices"][s:e], i] = a["weights"][s:e]
    op = plan.operator()
    M = np.zeros((n, m))
    at = 0
    for j in range(n):
        c = op["count"][j]
        M[j, op["lo"][j]:op["lo"][j] + c] = op["values"][at:at + c]
        at += c
    return A, M, op


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-k{c[2]}-b{c[5]}")
def test_operator_inverts_the_forward_model(case):
    m, n, kernel, taps, shift, blur = case
    r = make_request(m, n, float(n), shift, kernel, taps, blur=blur)
    A, M, op = _dense(_plan(r))
    assert op["dropped"] <= 2.0 ** -28
    eye = M @ A
    assert np.abs(eye - np.eye(n)).max() < 2e-7
    G = A.T @ A
    assert np.abs(G @ M - A.T).max() < 2e-7
    assert np.abs(M.sum(axis=1) - 1.0).max() < 2e-7


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: f"{c[0]}-{c[1]}")
def test_operator_equals_the_fp64_axis_solve(case, oracle):
    m, n, kernel, taps, shift, blur = case
    r = make_request(m, n, float(n), shift, kernel, taps, blur=blur)
    _, M, _ = _dense(_plan(r))
    y = np.random.default_rng(m + n).uniform(0, 1, (4, m))
    plan64 = oracle.build(fp64_request(r))
    ref = np.stack([oracle.axis_f64(plan64, row) for row in y])
    got = y @ M.T
    # F32 weights in the operator against FP64 weights in the oracle.
    assert np.abs(got - ref).max() < 1e-7


def test_band_is_contiguous_and_local():
    r = make_request(1920, 1280, 1280.0, 0.0, BICUBIC)
    _, M, op = _dense(_plan(r))
    assert op["max_band"] <= 56
    # Interior rows (the band is clipped to the plane near its edges).
    inner = np.arange(20, 1260)
    centre = (inner + 0.5) * 1.5 - 0.5
    mid = op["lo"][inner] + op["count"][inner] / 2.0
    assert np.abs(mid - centre).max() < 2.0

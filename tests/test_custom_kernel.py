"""Custom-kernel plans (dsmvc_hip_plan_build_custom) -- CPU.

Pinned two ways:
  * the bicubic polynomial handed in as a callback yields a plan bit-identical
    to the built-in bicubic plan (every array, and the reference's plan
    hashes), blur included: the custom path evaluates the same distances,
    normalises and coalesces the same way;
  * a kernel with no built-in counterpart (a Gaussian) yields the plan the
    reference engine's own custom-kernel path builds (oracle/_ref, reference
    src/axis_plan.cpp:535-599), bit for bit.
Failures keep the reference's messages.
"""
import math

import numpy as np
import pytest

import dsmvc_hip as dh
from support import BICUBIC, Request, complete_plan_hash, f32_plan_hash, make_request


def _bicubic(b, c):
    # The polynomial of the built-in kernel, in its evaluation order
    # (reference filter_weight, axis_plan.cpp:272-344).
    def w(x):
        if x < 1.0:
            return ((12.0 - 9.0 * b - 6.0 * c) * (x * x * x)
                    + (-18.0 + 12.0 * b + 6.0 * c) * (x * x) + (6.0 - 2.0 * b)) / 6.0
        if x < 2.0:
            return ((-b - 6.0 * c) * (x * x * x) + (6.0 * b + 30.0 * c) * (x * x)
                    + (-12.0 * b - 48.0 * c) * x + (8.0 * b + 24.0 * c)) / 6.0
        return 0.0
    return w


def _same_arrays(p, q):
    a, b = p.arrays(), q.arrays()
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)) for k in a)


def _as_request(r: dh.AxisRequest, kernel):
    return make_request(r.source_size, r.destination_size, r.active_length, r.shift, kernel,
                        r.taps, r.b, r.c, r.blur, r.border, r.f64_mode)


@pytest.mark.parametrize("sw,dw,shift,blur,b,c", [
    (1920, 1280, 0.0, 1.0, 0.0, 0.5), (960, 640, 1 / 12, 1.0, 1 / 3, 1 / 3),
    (1080, 980, 0.95, 1.3, 0.0, 0.75), (47, 39, -0.3125, 1.0, 0.0, 0.5)])
def test_bicubic_callback_equals_builtin_bicubic(sw, dw, shift, blur, b, c):
    builtin = dh.Plan(dh.AxisRequest(sw, dw, float(dw), shift, dh.BICUBIC, 3, b, c, blur))
    custom = dh.build_custom_plan(dh.AxisRequest(sw, dw, float(dw), shift, dh.CUSTOM, 2, b, c,
                                                 blur), _bicubic(b, c))
    assert custom.requires_float64 == builtin.requires_float64
    assert custom.normal_rcond == builtin.normal_rcond
    assert _same_arrays(custom, builtin)
    # The reference's plan hashes see the same plan (hashed as the bicubic request).
    req = _as_request(builtin.request, BICUBIC)
    arrays_c = _arrays_of(custom)
    arrays_b = _arrays_of(builtin)
    assert f32_plan_hash(req, arrays_c) == f32_plan_hash(req, arrays_b)
    assert complete_plan_hash(req, arrays_c) == complete_plan_hash(req, arrays_b)


def _arrays_of(plan):
    from support import PlanArrays
    a = plan.arrays()
    return PlanArrays(plan.source_size, plan.destination_size, plan.support, plan.half_bandwidth,
                      plan.active_length, plan.shift, plan.normal_rcond, plan.normal_inf_norm,
                      plan.requires_float64, **a)


def test_gaussian_kernel_equals_reference_custom_path(reference):
    def gauss(x):
        return math.exp(-x * x / 0.8)

    for sw, dw, taps, blur, mode in ((1920, 1280, 2, 1.0, 0), (540, 500, 3, 0.9, 0),
                                     (1080, 720, 2, 1.0, 2)):
        r = dh.AxisRequest(sw, dw, float(dw), 0.0, dh.CUSTOM, taps, 0.0, 0.5, blur,
                           dh.SYMMETRIC, mode)
        mine = dh.build_custom_plan(r, gauss)
        theirs = reference.build_custom(_as_request(r, dh.CUSTOM), gauss)
        ta = theirs.arrays()
        ma = mine.arrays()
        assert theirs.info.requires_float64 == int(mine.requires_float64)
        assert theirs.info.normal_rcond == mine.normal_rcond
        for k in ma:
            assert np.array_equal(ma[k].view(np.uint8), getattr(ta, k).view(np.uint8)), k


def test_custom_kernel_failures_keep_reference_messages():
    r = dh.AxisRequest(200, 150, 150.0, 0.0, dh.CUSTOM, 2)
    with pytest.raises(dh.DsmvcError, match="non-finite"):
        dh.build_custom_plan(r, lambda x: float("nan"))
    with pytest.raises(dh.DsmvcError, match="zero or non-finite weight sum"):
        dh.build_custom_plan(r, lambda x: 0.0)
    with pytest.raises(ValueError, match="callback is missing"):
        dh.Plan(r)
    with pytest.raises(ZeroDivisionError):
        dh.build_custom_plan(r, lambda x: 1.0 / 0.0)

"""CPU tests of the product library (no GPU): the C ABI loads and exports every
symbol include/dsmvc_hip.h declares, the library's FP64 planner reproduces the
oracle's plans bit for bit (and so the reference's golden hashes), and the
backend rules hold (explicit hip without a device fails, never falls back)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import dsmvc_hip as dh
from dsmvc_hip import _lib
from support import (Request, axis_fixtures, complete_plan_hash, f32_plan_hash, make_request,
                     PlanArrays)
from workloads import workload_plan_requests, integer_conversion as ref_integer_conversion

ROOT = Path(__file__).resolve().parents[1]


def _product_plan(req: Request) -> PlanArrays:
    r = dh.AxisRequest(req.source_size, req.destination_size, req.active_length, req.shift,
                       req.kernel, req.taps, req.b, req.c, req.blur, req.border, req.f64_mode)
    p = dh.Plan(r)
    a = p.arrays()
    return PlanArrays(p.source_size, p.destination_size, p.support, p.half_bandwidth,
                      p.active_length, p.shift, p.normal_rcond, p.normal_inf_norm,
                      p.requires_float64, a["offsets"], a["indices"], a["weights"],
                      a["lower_ld"], a["upper_l"], a["inverse_diagonal"], a["weights_f64"],
                      a["normal_bands_f64"], a["ldlt_bands_f64"], a["inverse_diagonal_f64"])


def test_header_symbols_exported():
    header = (ROOT / "include" / "dsmvc_hip.h").read_text()
    declared = set(re.findall(r"\b(dsmvc_hip_[a-z0-9_]+)\s*\(", header))
    assert declared == set(_lib.EXPORTS)
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("fx", axis_fixtures(), ids=lambda f: f.name)
def test_planner_reproduces_reference_goldens(fx):
    a = _product_plan(fx.request)
    assert a.half_bandwidth == fx.expected_half_bandwidth
    assert f32_plan_hash(fx.request, a) == fx.f32_plan_hash
    assert complete_plan_hash(fx.request, a) == fx.complete_plan_hash


def _equal(a: PlanArrays, b: PlanArrays):
    for name in ("support", "half_bandwidth", "normal_rcond", "normal_inf_norm",
                 "requires_float64"):
        assert getattr(a, name) == getattr(b, name), name
    for name in ("offsets", "indices", "weights", "lower_ld", "upper_l", "inverse_diagonal",
                 "weights_f64", "normal_bands_f64", "ldlt_bands_f64", "inverse_diagonal_f64"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), name


def _arrays_of(p: "dh.Plan") -> PlanArrays:
    a = p.arrays()
    return PlanArrays(p.source_size, p.destination_size, p.support, p.half_bandwidth,
                      p.active_length, p.shift, p.normal_rcond, p.normal_inf_norm,
                      p.requires_float64, a["offsets"], a["indices"], a["weights"],
                      a["lower_ld"], a["upper_l"], a["inverse_diagonal"], a["weights_f64"],
                      a["normal_bands_f64"], a["ldlt_bands_f64"], a["inverse_diagonal_f64"])


def _to_request(req: Request) -> "dh.AxisRequest":
    return dh.AxisRequest(req.source_size, req.destination_size, req.active_length, req.shift,
                          req.kernel, req.taps, req.b, req.c, req.blur, req.border, req.f64_mode)


def test_build_plans_on_threads_match_the_oracle(oracle):
    """dsmvc_hip_plan_build_many on 8 threads, from an empty cache: every plan
    equals the oracle's bit for bit, duplicates share one build, and the
    result is the same as one-at-a-time builds."""
    reqs = (workload_plan_requests("mitchell_21_heights")
            + workload_plan_requests("served_mix_sample")
            + workload_plan_requests("near_identity_blur09"))
    reqs = reqs + reqs[:6]                      # duplicates within one call
    dh.clear_planner_caches()
    plans = dh.build_plans([_to_request(r) for r in reqs], threads=8)
    stats = dh.planner_cache_stats()
    unique = len({tuple(getattr(r, f) for f, _ in r._fields_) for r in reqs})
    assert stats["plan_builds"] == unique
    for req, plan in zip(reqs, plans):
        _equal(_arrays_of(plan), oracle.build(req).arrays())
    dh.clear_planner_caches()
    for req, plan in zip(reqs[:8], plans[:8]):
        _equal(_arrays_of(plan), _product_plan(req))


def test_build_plans_validates_before_building():
    good = _to_request(make_request(1920, 1280, kernel=dh.BICUBIC))
    bad = dh.AxisRequest(1920, 1280, 1280.0, 0.0, 9)             # no such kernel
    dh.clear_planner_caches()
    with pytest.raises(ValueError, match="kernel is invalid"):
        dh.build_plans([good, bad, good], threads=2)
    assert dh.planner_cache_stats()["plan_builds"] == 0
    # A request that passes validation but cannot be planned: the error of the
    # lowest failing index, and no handles.
    empty = dh.AxisRequest(0, 10, 10.0, 0.0, dh.BICUBIC)
    with pytest.raises((ValueError, dh.DsmvcError)):
        dh.build_plans([good, empty], threads=2)
    assert dh.build_plans([], threads=2) == []


@pytest.mark.parametrize("workload", ["debilinear_1080p_720p", "yuv420p16_debicubic_1080p_720p",
                                      "near_identity_blur09", "mitchell_501_heights",
                                      "served_mix_sample"])
def test_planner_matches_oracle_at_workloads(oracle, workload):
    for req in workload_plan_requests(workload):
        _equal(_product_plan(req), oracle.build(req).arrays())


def test_planner_matches_oracle_edge_geometries(oracle):
    cases = [
        make_request(8, 1, kernel=dh.BICUBIC),              # one-sample destination, W = 0
        make_request(9, 3, kernel=dh.BICUBIC),              # W truncated to n - 1 = 2
        make_request(5, 5, kernel=dh.LANCZOS, taps=4),      # W = 4 (even)
        make_request(40, 17, 16.3, -2.7, dh.SPLINE36, border=dh.ZERO),
        make_request(40, 23, 22.0, 1.3, dh.SPLINE16, border=dh.MIRROR),
        make_request(33, 30, 30.0, 0.0, dh.LANCZOS, taps=3, blur=2.2),   # wide band
        make_request(100, 61, 60.5, 0.3, dh.BICUBIC, b=1 / 3, c=1 / 3, f64_mode=2),
    ]
    for req in cases:
        _equal(_product_plan(req), oracle.build(req).arrays())


def test_planner_errors_match_reference_kinds():
    with pytest.raises(ValueError, match="axis dimensions must be positive"):
        dh.Plan(dh.AxisRequest(0, 10, 10.0))
    with pytest.raises(ValueError, match="blur must be finite"):
        dh.Plan(dh.AxisRequest(20, 10, 10.0, blur=0.0))
    with pytest.raises(ValueError, match="padding mode is invalid"):
        dh.Plan(dh.AxisRequest(20, 10, 10.0, border=7))


def test_planner_cache_single_flight():
    dh.clear_planner_caches()
    r = dh.AxisRequest(96, 64, 64.0)
    a, b = dh.Plan(r), dh.Plan(r)
    stats = dh.planner_cache_stats()
    assert stats["plan_builds"] == 1 and stats["plan_hits"] == 1 and stats["plan_entries"] == 1
    dh.Plan(dh.AxisRequest(96, 64, 64.0, f64_mode=2))
    assert dh.planner_cache_stats()["plan_builds"] == 2


def test_geometry_reused_across_kernels_blur_and_precision(oracle):
    """The sampling geometry is keyed by sizes, active length, shift, support
    and border only (reference axis_plan.cpp:444-454, 520-533): lanczos3 and
    spline36 (support 3), bicubic at blur 1.0 and 0.9 (support ceil(1.8) = 2)
    and every f64mode share one geometry, and each plan built on a reused
    geometry equals the oracle's bit for bit."""
    dh.clear_planner_caches()
    shared = [make_request(1920, 1136, kernel=dh.LANCZOS, taps=3),
              make_request(1920, 1136, kernel=dh.SPLINE36),
              make_request(1920, 1136, kernel=dh.LANCZOS, taps=3, f64_mode=2),
              make_request(1920, 1136, kernel=dh.SPLINE36, f64_mode=1)]
    for req in shared:
        _equal(_product_plan(req), oracle.build(req).arrays())
    stats = dh.planner_cache_stats()
    assert stats["plan_builds"] == 4
    assert stats["geometry_builds"] == 1 and stats["geometry_hits"] == 3
    assert stats["geometry_entries"] == 1 and stats["geometry_resident_bytes"] > 0
    blur = [make_request(1920, 1136, kernel=dh.BICUBIC),
            make_request(1920, 1136, kernel=dh.BICUBIC, blur=0.9),
            make_request(1920, 1136, kernel=dh.BICUBIC, b=1 / 3, c=1 / 3, blur=0.95)]
    for req in blur:
        _equal(_product_plan(req), oracle.build(req).arrays())
    stats = dh.planner_cache_stats()
    assert stats["geometry_builds"] == 2 and stats["geometry_hits"] == 5
    # A different shift or border is a different geometry.
    for req in (make_request(1920, 1136, 1136.0, 0.25, dh.BICUBIC),
                make_request(1920, 1136, kernel=dh.BICUBIC, border=dh.ZERO)):
        _equal(_product_plan(req), oracle.build(req).arrays())
    assert dh.planner_cache_stats()["geometry_builds"] == 4
    dh.clear_planner_caches()
    stats = dh.planner_cache_stats()
    assert stats["geometry_entries"] == 0 and stats["geometry_builds"] == 0


def test_padding_index_table():
    expect = [2, 1, 0, 1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1]
    assert [dh.padding_index(-8 + k, 4, dh.REFLECT101) for k in range(20)] == expect
    with pytest.raises(ValueError):
        dh.padding_index(0, 4, 9)


def test_backend_names_and_no_fallback():
    assert dh.parse_backend("hip") == "hip"
    assert dh.parse_backend("auto") == "auto"
    with pytest.raises(ValueError):
        dh.parse_backend("opencl")
    if dh.device_count() == 0:
        with pytest.raises(dh.DsmvcError, match="unavailable"):
            dh.resolve_backend("hip")
        with pytest.raises(dh.DsmvcError, match="unavailable"):
            dh.Executor()
    with pytest.raises(dh.DsmvcError):
        dh.resolve_backend("cuda")


def test_integer_conversion_constants():
    for bits in (8, 10, 16):
        for chroma in (False, True):
            for full in (False, True):
                a = dh.integer_conversion(bits, chroma, full)
                b = ref_integer_conversion(bits, chroma, full)
                assert (a.input_offset, a.input_scale, a.output_scale, a.output_offset,
                        a.output_maximum) == (b.input_offset, b.input_scale, b.output_scale,
                                              b.output_offset, b.output_maximum)
    c = dh.integer_conversion(16, False, False)
    assert (c.input_offset, c.output_scale) == (4096.0, 56064.0)
    c = dh.integer_conversion(16, True, False)
    assert (c.input_offset, c.output_scale) == (32768.0, 57344.0)

"""Pins the oracle (oracle/dsmvc_oracle.c) before anything is checked against it.

Checks, each chosen so that a plausible restatement mistake fails one of them:
  * the reference's 7 axis fixtures x 4 golden FNV hashes
    (tests/numerical_conformance.hpp:126-164) -- plan arrays and both solvers;
  * the mixed safe/risky 2D golden a9dcffe6649bd96e (numerical_conformance.hpp:172-176,
    numerical_contract_tests.cpp:645-689) -- the Float64 2D route;
  * the QR anchor (numerical_contract_tests.cpp:697-726) and the 60-digit
    conditioned-bilinear anchor (engine_tests.cpp:992-1031);
  * the padding-index tables (engine_tests.cpp:1334-1387);
  * identity bilinear (engine_tests.cpp:954-969);
  * the long-double normal-matrix audit (numerical_conformance.hpp:274-328);
  * dense float64 least-squares solves (numpy) on small systems and the
    forward-model round trip -- mathematics, not the reference;
  * the reference engine compiled from its own sources (oracle/_ref), bit for bit,
    at every workload geometry (plans, scalar 2D F32, integer, Float64 routes).
"""
import numpy as np
import pytest

from support import (BICUBIC, BILINEAR, F64_ALWAYS, F64_AUTO, F64_NEVER, LANCZOS, MIRROR,
                     MIXED_INPUT_SEED, MIXED_OUTPUT_HASH, REFLECT101, REPEAT, SPLINE16,
                     SPLINE36, SPLINE64, SYMMETRIC, ZERO, Conversion, axis_fixtures,
                     complete_plan_hash, conditioned_lanczos2_request, f32_plan_hash,
                     make_normal_input, make_request, mixed_horizontal_request, output_hash,
                     ulp_distance)
from workloads import workload_plan_requests, integer_conversion


@pytest.mark.parametrize("fx", axis_fixtures(), ids=lambda f: f.name)
def test_fixture_goldens(oracle, fx):
    plan = oracle.build(fx.request)
    a = plan.arrays()
    assert a.half_bandwidth == fx.expected_half_bandwidth
    assert f32_plan_hash(fx.request, a) == fx.f32_plan_hash
    assert complete_plan_hash(fx.request, a) == fx.complete_plan_hash
    x = make_normal_input(a.source_size, fx.input_seed)
    production = oracle.axis_f32(plan, x)
    assert output_hash(production) == fx.production_output_hash
    if a.requires_float64:
        ordered = oracle.axis_f64_ordered(plan, x.astype(np.float64)).astype(np.float32)
        assert int(ulp_distance(ordered, production).max()) <= 1
    else:
        ordered = oracle.axis_f32_ordered(plan, x)
    assert output_hash(ordered) == fx.ordered_output_hash


def test_fixture_rcond_routes(oracle):
    fx = axis_fixtures()
    for i, f in enumerate(fx):
        rcond = oracle.build(f.request).arrays().normal_rcond
        if i in (4, 6):
            assert rcond < 1e-4, f.name
        else:
            assert rcond >= 1e-4, f.name


def _mixed_2d(oracle, h, v, seed):
    hs, vs = h.c.source_size, v.c.source_size
    packed = make_normal_input(vs * hs, seed).reshape(vs, hs)
    x = np.full((vs, hs + 2), 7.0, np.float32)
    x[:, :hs] = packed
    inter = np.stack([oracle.axis_f64_ordered(h, row.astype(np.float64)) for row in packed])
    ordered = np.stack([oracle.axis_f64_ordered(v, inter[:, c])
                        for c in range(h.c.destination_size)], axis=1).astype(np.float32)
    production = oracle.inverse_2d(h, v, x[:, :hs])
    return ordered, production


def test_mixed_safe_risky_2d_golden(oracle):
    safe = oracle.build(mixed_horizontal_request())
    risky = oracle.build(conditioned_lanczos2_request(F64_AUTO))
    assert not safe.c.requires_float64 and risky.c.requires_float64
    ordered, production = _mixed_2d(oracle, safe, risky, MIXED_INPUT_SEED)
    assert output_hash(ordered.ravel()) == MIXED_OUTPUT_HASH
    assert output_hash(production.ravel()) == MIXED_OUTPUT_HASH


def test_qr_anchor(oracle):
    plan = oracle.build(conditioned_lanczos2_request(F64_AUTO))
    state = 0x9781A5C3
    x = np.empty(plan.c.source_size)
    for i in range(x.size):
        state ^= (state << 13) & 0xFFFFFFFF
        state ^= state >> 17
        state ^= (state << 5) & 0xFFFFFFFF
        x[i] = float(np.float32(state & 0xFFFF) / np.float32(65535.0))
    out = oracle.axis_f64_ordered(plan, x)
    first, last = 74.22802437585693, -119.91860242539072
    error = max(abs(out[0] - first), abs(out[-1] - last))
    bound = 4.0 * np.finfo(np.float64).eps / plan.c.normal_rcond * max(1.0, abs(first), abs(last))
    assert error <= bound


def test_conditioned_bilinear_anchor(oracle):
    req = make_request(1920, 1734, 1732.0888888888887, 0.9555555555556339, BILINEAR)
    plan = oracle.build(req)
    assert plan.c.requires_float64
    assert 2.0e-6 < plan.c.normal_rcond < 2.5e-6
    state = 0x9743B111
    x = np.empty(1920, np.float32)
    for i in range(x.size):
        state ^= (state << 13) & 0xFFFFFFFF
        state ^= state >> 17
        state ^= (state << 5) & 0xFFFFFFFF
        x[i] = np.float32(state & 0xFFFF) / np.float32(65535.0)
    robust = oracle.axis_f32(plan, x)
    error = max(abs(float(robust[0]) - 64.704580093370353),
                abs(float(robust[-1]) - 6.4342414526327527))
    assert error < 3.0e-6


def test_padding_tables(oracle):
    first = -8
    expected = {
        ZERO: [-1] * 8 + [0, 1, 2, 3] + [-1] * 8,
        REPEAT: [0] * 8 + [0, 1, 2, 3] + [3] * 8,
        REFLECT101: [2, 1, 0, 1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1],
        SYMMETRIC: [0, 1, 2, 3, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, 0, 1, 2, 3],
        MIRROR: [-1, -1, -1, -1, 3, 2, 1, 0, 0, 1, 2, 3, 3, 2, 1, 0, -1, -1, -1, -1],
    }
    for mode, table in expected.items():
        got = [oracle.padding_index(first + k, 4, mode) for k in range(20)]
        assert got == table, mode
    for mode in (REPEAT, REFLECT101, SYMMETRIC):
        assert oracle.padding_index(-1000000000000, 1, mode) == 0
        assert oracle.padding_index(1000000000000, 1, mode) == 0
    assert oracle.padding_index(0, 4, 255) == -2


def test_identity_bilinear(oracle):
    plan = oracle.build(make_request(32, 32, 32.0, 0.0, BILINEAR))
    x = (np.arange(32, dtype=np.float32) / np.float32(31.0)).astype(np.float32)
    out = oracle.axis_f32(plan, x)
    assert np.abs(out - x).max() < 1e-5


def _dense_A(a):
    """Dense A (source x destination) from the CSR of A^T (float64 weights)."""
    A = np.zeros((a.source_size, a.destination_size))
    w = a.weights_f64 if a.requires_float64 else a.weights.astype(np.float64)
    for i in range(a.destination_size):
        for k in range(a.offsets[i], a.offsets[i + 1]):
            A[a.indices[k], i] = w[k]
    return A


@pytest.mark.parametrize("kernel,taps,border", [
    (BILINEAR, 3, SYMMETRIC), (BICUBIC, 3, REFLECT101), (LANCZOS, 3, REPEAT),
    (SPLINE16, 3, ZERO), (SPLINE36, 3, SYMMETRIC), (SPLINE64, 3, MIRROR)])
def test_normal_matrix_and_dense_solve(oracle, kernel, taps, border):
    req = make_request(53, 37, 36.5, 0.3, kernel, taps, border=border, f64_mode=F64_ALWAYS)
    a = oracle.build(req).arrays()
    n, w = a.destination_size, a.half_bandwidth
    A = _dense_A(a)
    # Row sums of A are 1 (weights normalised by their tap sum) unless all taps fell
    # outside the zero/mirror border.
    sums = A.sum(axis=1)
    assert np.all((np.abs(sums - 1) < 1e-12) | (border in (ZERO, MIRROR)))
    # Long-double re-formation of G = A^T A against the retained bands.
    Al = A.astype(np.longdouble)
    G = Al.T @ Al
    for d in range(w + 1):
        band = a.normal_bands_f64[d * n:(d + 1) * n][: n - d]
        ref = np.array([G[i, i + d] for i in range(n - d)], dtype=np.longdouble)
        tol = 128 * np.finfo(np.float64).eps * max(1.0, float(np.abs(G).max()))
        assert float(np.abs(ref - band.astype(np.longdouble)).max()) <= tol
    # Entries outside the band are exactly zero.
    Gd = A.T @ A
    for i in range(n):
        for j in range(n):
            if abs(i - j) > w:
                assert Gd[i, j] == 0.0
    # The F64 route solves the normal equations.
    rng = np.random.default_rng(5)
    y = rng.uniform(0, 1, a.source_size)
    x = oracle.axis_f64(oracle.build(req), y)
    expected = np.linalg.solve(Gd + np.eye(n) * np.finfo(float).eps, A.T @ y)
    assert np.abs(x - expected).max() < 1e-9 * max(1.0, np.abs(expected).max())


def test_forward_model_round_trip(oracle):
    """y = A x_true => the descale recovers x_true (a well-conditioned downscale)."""
    req = make_request(96, 64, kernel=BICUBIC, f64_mode=F64_NEVER)
    plan = oracle.build(req)
    A = _dense_A(plan.arrays())
    x_true = np.random.default_rng(1).uniform(0, 1, 64)
    y = (A @ x_true).astype(np.float32)
    x = oracle.axis_f32(plan, y)
    assert np.abs(x - x_true).max() < 2e-5


def test_invalid_requests(oracle):
    with pytest.raises(ValueError):
        oracle.build(make_request(0, 10))
    with pytest.raises(ValueError):
        oracle.build(make_request(20, 10, blur=0.0))
    with pytest.raises(ValueError):
        oracle.build(make_request(20, 10, active_length=-1.0))
    with pytest.raises(ValueError):
        oracle.build(make_request(20, 10, border=9))


# ---------------------------------------------------------------------------
# The oracle against the reference engine built from its own sources
# ---------------------------------------------------------------------------
pytestmark_ref = pytest.mark.reference


def _plan_equal(a, b):
    assert a.support == b.support and a.half_bandwidth == b.half_bandwidth
    assert a.normal_rcond == b.normal_rcond
    assert a.normal_inf_norm == b.normal_inf_norm
    assert a.requires_float64 == b.requires_float64
    for name in ("offsets", "indices", "weights", "lower_ld", "upper_l", "inverse_diagonal",
                 "weights_f64", "normal_bands_f64", "ldlt_bands_f64", "inverse_diagonal_f64"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.shape == y.shape, name
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name


@pytest.mark.reference
@pytest.mark.parametrize("workload", ["debilinear_1080p_720p", "yuv420p16_debicubic_1080p_720p",
                                      "near_identity_blur09", "mitchell_21_heights",
                                      "served_mix_sample"])
def test_plans_match_reference_engine(oracle, reference, workload):
    for req in workload_plan_requests(workload):
        _plan_equal(oracle.build(req).arrays(), reference.build(req).arrays())


@pytest.mark.reference
def test_scalar_2d_matches_reference_engine(oracle, reference):
    rng = np.random.default_rng(11)
    for kernel in (BILINEAR, BICUBIC, LANCZOS, SPLINE36):
        h = make_request(480, 320, kernel=kernel)
        v = make_request(270, 180, kernel=kernel)
        x = rng.uniform(0, 1, (270, 480)).astype(np.float32)
        o = oracle.inverse_2d(oracle.build(h), oracle.build(v), x)
        r = reference.inverse_2d(reference.build(h), reference.build(v), x)
        assert np.array_equal(o.view(np.uint32), r.view(np.uint32))


@pytest.mark.reference
def test_float64_2d_matches_reference_engine(oracle, reference):
    rng = np.random.default_rng(12)
    h = make_request(400, 397, kernel=LANCZOS, taps=4, blur=0.9)
    v = make_request(240, 120, kernel=SPLINE64, blur=0.9, f64_mode=F64_ALWAYS)
    x = rng.uniform(0, 1, (240, 400)).astype(np.float32)
    o = oracle.inverse_2d(oracle.build(h), oracle.build(v), x)
    r = reference.inverse_2d(reference.build(h), reference.build(v), x)
    assert np.array_equal(o.view(np.uint32), r.view(np.uint32))


@pytest.mark.reference
@pytest.mark.parametrize("dtype,bits", [(np.uint16, 16), (np.uint8, 8), (np.uint16, 10)])
def test_integer_2d_matches_reference_engine(oracle, reference, dtype, bits):
    rng = np.random.default_rng(bits)
    h = make_request(480, 320, kernel=BICUBIC)
    v = make_request(270, 180, kernel=BICUBIC)
    x = rng.integers(0, 1 << bits, (270, 480)).astype(dtype)
    conv = integer_conversion(bits, chroma=False, full_range=False)
    o = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x, conv)
    rh, rv = reference.build(h), reference.build(v)
    r = reference.inverse_2d_int(rh, rv, x, conv)
    rs = reference.inverse_2d_int(rh, rv, x, conv, streamed=True)
    assert np.array_equal(o, r) and np.array_equal(o, rs)


@pytest.mark.reference
def test_integer_2d_f64_matches_reference_engine(oracle, reference):
    rng = np.random.default_rng(3)
    h = make_request(300, 200, kernel=BICUBIC, f64_mode=F64_ALWAYS)
    v = make_request(160, 100, kernel=BICUBIC)
    x = rng.integers(0, 65536, (160, 300)).astype(np.uint16)
    conv = integer_conversion(16, chroma=True, full_range=False)
    o = oracle.inverse_2d_int(oracle.build(h), oracle.build(v), x, conv)
    r = reference.inverse_2d_int(reference.build(h), reference.build(v), x, conv)
    assert np.array_equal(o, r)


@pytest.mark.reference
def test_full_size_1080p_scalar_matches_reference_engine(oracle, reference):
    """Config-1/2 geometry at full size: scalar 2D and u16 integer, bit for bit."""
    rng = np.random.default_rng(1)
    reqs = workload_plan_requests("yuv420p16_debicubic_1080p_720p")
    h, v = reqs[0], reqs[1]
    x = rng.uniform(0, 1, (1080, 1920)).astype(np.float32)
    o = oracle.inverse_2d(oracle.build(h), oracle.build(v), x)
    r = reference.inverse_2d(reference.build(h), reference.build(v), x)
    assert np.array_equal(o.view(np.uint32), r.view(np.uint32))
    xi = rng.integers(0, 65536, (1080, 1920)).astype(np.uint16)
    conv = integer_conversion(16, chroma=False, full_range=False)
    assert np.array_equal(oracle.inverse_2d_int(oracle.build(h), oracle.build(v), xi, conv),
                          reference.inverse_2d_int(reference.build(h), reference.build(v), xi,
                                                   conv))

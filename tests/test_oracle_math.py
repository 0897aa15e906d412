"""Pins the oracle's scalar building blocks against published answers and
libm, and checks the FP32 rewrites the GPU kernels rely on."""
import math

import numpy as np
import pytest

# Random123 known-answer vectors for Philox4x32-10 (ctr, key) -> out.
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle_lib, ctr, key, out):
    assert [int(x) for x in oracle_lib.philox(ctr, key)] == out


def _ulp_err(got: float, want: float) -> float:
    if want == 0.0:
        return abs(got) / np.finfo(np.float32).tiny
    ulp = np.spacing(np.float32(abs(want)))
    return abs(got - want) / float(ulp)


@pytest.mark.parametrize("op,fn,lo,hi,maxulp", [
    (2, math.sin, 0.0, 8.0, 2.5),
    (3, math.cos, 0.0, 8.0, 2.5),
    (5, math.exp, -80.0, 80.0, 2.5),
    (6, math.log, 1e-30, 1e30, 2.5),
])
def test_transcendentals_vs_libm(oracle_lib, op, fn, lo, hi, maxulp):
    rng = np.random.default_rng(op)
    if op == 6:
        xs = np.exp(rng.uniform(math.log(lo), math.log(hi), 4000)).astype(np.float32)
    else:
        xs = rng.uniform(lo, hi, 4000).astype(np.float32)
    worst = 0.0
    for x in xs:
        want = fn(float(x))
        got = oracle_lib.math(op, float(x))
        if op in (2, 3) and abs(want) < 1e-3:
            assert abs(got - want) < 4e-7
            continue
        worst = max(worst, _ulp_err(got, want))
    assert worst <= maxulp, worst


def test_pow_vs_libm(oracle_lib):
    """pow(x, y) = exp(y log x) for the SVGF normal weight: relative error bounded by
    the rounding of y log x, |y log x| * 2^-24, plus a few ulp of exp."""
    rng = np.random.default_rng(10)
    xs = rng.uniform(0.0, 1.0, 4000).astype(np.float32)
    ys = rng.uniform(0.0, 300.0, 4000).astype(np.float32)
    for x, y in zip(xs, ys):
        want = math.pow(float(x), float(y))
        if want < 1e-30:
            continue
        got = oracle_lib.math(10, float(x), float(y))
        tol = 3e-7 + abs(float(y) * math.log(float(x))) * 1.2e-7
        assert abs(got - want) <= tol * want, (x, y, got, want)
    assert oracle_lib.math(10, 0.0, 2.0) == 0.0
    assert oracle_lib.math(10, 0.0, 0.0) == 1.0
    assert oracle_lib.math(10, 1.0, 256.0) == 1.0
    assert oracle_lib.math(10, 0.5, 1.0) == 0.5


def test_acos_and_atan2_vs_libm(oracle_lib):
    rng = np.random.default_rng(7)
    for x in rng.uniform(-1, 1, 4000).astype(np.float32):
        assert abs(oracle_lib.math(4, float(x)) - math.acos(float(x))) <= 4e-7
    for y, x in rng.standard_normal((4000, 2)).astype(np.float32):
        assert abs(oracle_lib.math(7, float(y), float(x)) - math.atan2(float(y), float(x))) <= 6e-7
    # special cases of the quadrant logic
    assert oracle_lib.math(7, 0.0, -1.0) == pytest.approx(math.pi, abs=1e-6)
    assert oracle_lib.math(7, -1.0, 0.0) == pytest.approx(-math.pi / 2, abs=1e-6)
    assert oracle_lib.math(7, 0.0, 0.0) == 0.0
    assert math.isnan(oracle_lib.math(4, 1.5))


def _float_neighbourhood(center: float, n: int = 4000) -> np.ndarray:
    c = np.float32(center)
    steps = np.arange(-n, n + 1)
    out = np.empty(steps.size, np.float32)
    cur_up, cur_dn = c, c
    out[n] = c
    for k in range(1, n + 1):
        cur_up = np.nextafter(cur_up, np.float32(np.inf))
        cur_dn = np.nextafter(cur_dn, np.float32(-np.inf))
        out[n + k] = cur_up
        out[n - k] = cur_dn
    return out


def test_slab_slack_is_exact_in_fp32():
    """(double)x + 1e-2 >= 0  <=>  x >= -0.01f for every float x near the threshold."""
    xs = _float_neighbourhood(-0.01)
    lhs = xs.astype(np.float64) + 1e-2 >= 0
    rhs = xs >= np.float32(-0.01)
    assert np.array_equal(lhs, rhs)


def test_shadow_tolerance_is_exact_in_fp32():
    """(double)x > 1e-1  <=>  x >= 0.1f for every float x near the threshold."""
    xs = _float_neighbourhood(0.1)
    assert np.array_equal(xs.astype(np.float64) > 1e-1, xs >= np.float32(0.1))


def test_reciprocal_through_double_is_correctly_rounded():
    """(float)(1.0 / (double)det) equals the FP32 quotient 1.0f / det (double
    rounding is innocuous for 53 >= 2*24 + 2)."""
    rng = np.random.default_rng(3)
    det = np.abs(rng.standard_normal(200000)).astype(np.float32) * np.float32(10.0) ** rng.integers(-20, 20, 200000).astype(np.float32)
    det = det[(det > 0) & np.isfinite(det)]
    a = (1.0 / det.astype(np.float64)).astype(np.float32)
    b = np.float32(1.0) / det
    assert np.array_equal(a.view(np.int32), b.view(np.int32))

"""Exactness facts the SVGF kernels rely on, checked on the CPU:

* the tap-distance table in svgf.h holds the correctly rounded float32 square
  roots of 0..18 (numpy's float32 sqrt is correctly rounded), so reading a
  tap's distance from it equals taking sqrtf of the squared offset;
* the variance pass's luminance guide replaces sqrt(max(a, b, 1e-4)) by
  max(sqrt(max(a, 1e-4)), sqrt(max(b, 1e-4))): equal floats because a
  correctly rounded square root is monotone -- checked on random, tied,
  denormal, huge and NaN operands with fmaxf's NaN rule (the non-NaN operand
  wins)."""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SVGF_H = os.path.join(HERE, "..", "cudaraytracer_amd", "csrc", "svgf.h")


def _tap_table():
    src = open(SVGF_H).read()
    m = re.search(r"kTapDist\[19\]\s*=\s*\{([^}]*)\}", src)
    assert m, "kTapDist table not found"
    vals = [v.strip().rstrip("f") for v in m.group(1).split(",") if v.strip()]
    return np.array([float.fromhex(v) for v in vals], np.float32)


def test_tap_distance_table_is_correctly_rounded_sqrt():
    t = _tap_table()
    assert t.size == 19
    want = np.sqrt(np.arange(19, dtype=np.float32))
    assert np.array_equal(t.view(np.int32), want.view(np.int32))
    # every squared offset of the 5x5 and 7x7 filters indexes the table
    sq = {i * i + j * j for i in range(-3, 4) for j in range(-3, 4)}
    assert max(sq) < t.size


def _fmaxf(a, b):
    """C fmaxf: the non-NaN operand if one is NaN."""
    return np.where(np.isnan(a), b, np.where(np.isnan(b), a, np.maximum(a, b))).astype(np.float32)


def test_sqrt_of_max_equals_max_of_sqrts():
    rng = np.random.default_rng(7)
    n = 200000
    a = (rng.random(n) ** 6 * 10.0 ** rng.integers(-8, 8, n)).astype(np.float32)
    b = (rng.random(n) ** 6 * 10.0 ** rng.integers(-8, 8, n)).astype(np.float32)
    b[::7] = a[::7]                                  # ties
    b[1::11] = np.nextafter(a[1::11], np.float32(np.inf))   # neighbours
    special = np.array([0.0, 1e-4, np.nextafter(np.float32(1e-4), 0), 1e-45, 3e38, np.nan], np.float32)
    a = np.concatenate([a, np.repeat(special, special.size)])
    b = np.concatenate([b, np.tile(special, special.size)])
    eps = np.float32(1e-4)
    with np.errstate(invalid="ignore"):
        lhs = np.sqrt(_fmaxf(_fmaxf(a, b), eps))
        rhs = _fmaxf(np.sqrt(_fmaxf(a, eps)), np.sqrt(_fmaxf(b, eps)))
    assert np.array_equal(lhs.view(np.int32), rhs.view(np.int32))

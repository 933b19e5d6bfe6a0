"""The rasterizer's barycentric quotient.

The block walk computes float(E) / area as RN32(RN64(double(float(E)) *
RN64(1 / area))) (``quot_f32`` in ``vct_amd/csrc/raster_common.h``): one FP64
multiply by the triangle's reciprocal instead of an IEEE float division per
edge and pixel.  The device comment proves the two are equal for a >= 0,
b > 0 floats with a normal quotient: a quotient of two 24-bit floats is never a
float rounding midpoint and lies at least 2^-50 (relatively) away from every
one, while the FP64 product errs by less than 2^-52.  These checks run the same
arithmetic in numpy (IEEE binary32 / binary64, round to nearest even) over the
operand ranges the rasterizer produces: edge values 0 .. 2^41 and areas up to
2^41 in 1/256-pixel units, with areas and edge values at and around powers of
two, where the float spacing changes.
"""
import numpy as np


def quot_f32(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    rcp = 1.0 / b.astype(np.float64)
    return (a.astype(np.float64) * rcp).astype(np.float32)


def _pairs(rng, n, kind):
    bits = rng.integers(2, 42, n)
    ai = rng.integers(1, np.int64(1) << 62, n, dtype=np.int64) >> (62 - bits)
    ai = np.maximum(ai, 1)
    if kind == "pow2":
        k = rng.integers(1, 41, n)
        ai = np.maximum((np.int64(1) << k) + rng.integers(-3, 4, n), 1)
    area = ai.astype(np.float32)
    if kind == "pow2_e":
        E = (np.int64(1) << rng.integers(0, 41, n)) + rng.integers(-2, 3, n)
        E = np.clip(E, 0, ai)
    else:
        E = (rng.random(n) * ai * 1.2).astype(np.int64)
    return E.astype(np.float32), area


def test_quotient_equals_ieee_division():
    rng = np.random.default_rng(20261203)
    for kind in ("uniform", "pow2", "pow2_e"):
        for _ in range(3):
            a, b = _pairs(rng, 2_000_000, kind)
            want = a / b  # numpy float32 division: correctly rounded
            got = quot_f32(a, b)
            bad = np.flatnonzero(want.view(np.uint32) != got.view(np.uint32))
            assert bad.size == 0, (kind, a[bad[:3]], b[bad[:3]])


def test_quotient_every_area_significand():
    # every float significand of the area in [2^20, 2^21), against edge values
    # spread over the same binade and the one below
    b = np.arange(1 << 23, dtype=np.uint32) | np.uint32(0x49800000)  # 2^20 .. 2^21
    b = b.view(np.float32)
    rng = np.random.default_rng(7)
    for _ in range(4):
        a = np.floor(rng.random(b.size) * b.astype(np.float64) * 1.5).astype(np.float32)
        want = a / b
        got = quot_f32(a, b)
        assert np.array_equal(want.view(np.uint32), got.view(np.uint32))


def test_a_float_reciprocal_is_not_enough():
    # the same product with a float reciprocal (what a fast division would
    # do) differs from the division: the FP64 reciprocal is what the proof needs
    rng = np.random.default_rng(11)
    a, b = _pairs(rng, 1_000_000, "uniform")
    want = a / b
    fast = a * (np.float32(1.0) / b)
    assert np.count_nonzero(want.view(np.uint32) != fast.view(np.uint32)) > 1000

"""Pins the oracle's storage conversions: IEEE binary16 against numpy, OCP FP8
E4M3 against a brute-force nearest-code search over its own 256-code table,
and GLSL packUnorm4x8 against its closed form."""
import ctypes as C

import numpy as np
import pytest


def test_half_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.uniform(-2, 2, 4000), rng.uniform(0, 1, 4000), 10.0 ** rng.uniform(-9, 5, 2000),
        [0.0, -0.0, 1e-8, 5.96e-8, 6.1e-5, 65504.0, 65519.99, 65520.0, 1e6, -1e6, 0.5, 1.0 / 3.0],
    ]).astype(np.float32)
    L = oracle.lib()
    for v in vals:
        h = L.or_float_to_half(float(v))
        ref = np.float32(v).astype(np.float16).view(np.uint16)
        assert h == ref, (v, hex(h), hex(int(ref)))
        if np.isfinite(np.float16(v)):
            assert L.or_half_to_float(h) == np.float32(np.float16(v))


def _e4m3_table():
    """All finite non-negative E4M3 values (OCP): bias 7, 0x7f is NaN."""
    vals = []
    for code in range(0x7f):
        e, m = code >> 3, code & 7
        vals.append(m * 2.0 ** -9 if e == 0 else (1 + m / 8) * 2.0 ** (e - 7))
    return np.array(vals)


def test_fp8_decode_table(oracle):
    table = _e4m3_table()
    assert table[-1] == 448.0
    for code in range(0x7f):
        assert oracle.fp8_decode(code) == table[code]
        assert oracle.fp8_decode(code | 0x80) == -table[code]
    assert np.isnan(oracle.fp8_decode(0x7f))


def test_fp8_encode_is_nearest_even_and_saturating(oracle):
    table = _e4m3_table()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0, 500, 3000), 10.0 ** rng.uniform(-4, 2.7, 3000),
                         (table[:-1] + table[1:]) / 2,  # exact midpoints: ties to the even code
                         table]).astype(np.float32)
    for x in xs:
        xv = float(np.float32(x))
        if xv >= 448.0:
            expect = 0x7e
        else:
            d = np.abs(table - xv)
            best = np.flatnonzero(d == d.min())
            expect = int(best[0]) if len(best) == 1 else int([b for b in best if b % 2 == 0][0])
        assert oracle.fp8_encode(xv) == expect, (xv, oracle.fp8_encode(xv), expect)
        assert oracle.fp8_encode(-xv) == (expect | 0x80)


def test_pack_unorm_closed_form(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for _ in range(2000):
        c = rng.uniform(-0.2, 1.2, 4).astype(np.float32)
        got = L.or_pack_unorm4x8(c.ctypes.data)
        q = np.rint(np.clip(c, 0, 1).astype(np.float32) * np.float32(255)).astype(np.uint32)
        assert got == int(q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24))
    # k/255 round-trips to k: the voxelize -> transferVoxels rewrite keeps albedo bytes
    for k in range(256):
        c = np.array([k / 255.0, 0, 0, 0], np.float32)
        assert L.or_pack_unorm4x8(c.ctypes.data) == k

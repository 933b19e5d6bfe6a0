"""Pins the oracle's warp texture (or_warp_occupancy, or_warpmap,
or_warp_sample: src/Application.cpp:235-575, generateWarpmap.frag) by
properties and independent restatements.

  occupancy     the 32^3 occupancy grid is exactly the occupancy of a plain
                voxelization at dim 32 (the same rasterizer and coverage)
  weights       every row's cell resolutions sum to 32 (so the warp maps
                [0, 1] onto [0, 1]); low/high follow the clamp rule of
                Application.cpp:346-370
  texel values  an independent restatement: warped centre = (sum of the
                resolutions of the cells before + half its own) / 32, from a
                numpy cumulative sum, quantized to unorm16
  identity      all-empty and all-occupied grids, warpTextureLinear and
                per-axis switches give back the texel centres
  sampling      LINEAR + CLAMP_TO_EDGE: exact texel values at centres, linear
                in between, clamped outside
"""
import numpy as np
import pytest

import vct_amd
from oracle import binding as ob


def _q16(v):
    return np.floor(np.clip(np.asarray(v, np.float64), 0, 1) * 65535 + 0.5) / 65535


def _weights(n, high=2.0, low=0.5, fp16=True):
    if n in (0, 32):
        lo = hi = 1.0
    else:
        hi = np.float32(high)
        lo = np.float32((np.float32(32) - hi * np.float32(n)) / np.float32(32 - n))
        if lo < np.float32(low):
            lo = np.float32(low)
            hi = np.float32((np.float32(32) - lo * np.float32(32 - n)) / np.float32(n))
    if fp16:
        lo, hi = float(np.float16(lo)), float(np.float16(hi))
    return float(lo), float(hi)


def _expect_x(occ, fp16=True):
    """warped x centre of every cell, by a cumulative sum of resolutions per x-row"""
    out = np.zeros(occ.shape)
    for z in range(32):
        for y in range(32):
            row = occ[z, y] > 0
            lo, hi = _weights(int(row.sum()), fp16=fp16)
            res = np.where(row, hi, lo)
            out[z, y] = (np.cumsum(res) - res / 2) / 32
    return out


def test_occupancy_is_the_dim32_voxel_occupancy(oracle):
    app = vct_amd.App(1, device=-1)
    fs = app.frame_state(0)
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    occ = ob.warp_occupancy(sc, fs["vox"])
    alb, _, _ = ob.voxelize(sc, fs["lights"], [None], fs["vox"], 32, lighting=0)
    assert np.array_equal(occ > 0, (alb >> 24) > 0)
    assert 0 < (occ > 0).sum() < 32 ** 3


@pytest.mark.parametrize("fill", [0, 1])
def test_uniform_grid_is_the_identity(oracle, fill):
    wm = ob.warpmap(np.full((32, 32, 32), fill, np.uint32))
    c = (np.arange(32) + 0.5) / 32
    assert np.array_equal(wm[..., 0], np.broadcast_to(_q16(c)[None, None, :], (32, 32, 32)).astype(np.float32))
    assert np.array_equal(wm[..., 1], np.broadcast_to(_q16(c)[None, :, None], (32, 32, 32)).astype(np.float32))
    assert np.array_equal(wm[..., 2], np.broadcast_to(_q16(c)[:, None, None], (32, 32, 32)).astype(np.float32))


@pytest.mark.parametrize("seed,frac", [(1, 0.1), (2, 0.3), (3, 0.7)])
def test_texels_match_cumulative_resolution(oracle, seed, frac):
    occ = (np.random.default_rng(seed).random((32, 32, 32)) < frac).astype(np.uint32)
    wm = ob.warpmap(occ)
    assert np.abs(wm[..., 0] - _q16(_expect_x(occ))).max() <= 1.5 / 65535
    # y and z: the same rule on the transposed grids
    assert np.abs(wm[..., 1] - np.transpose(_q16(_expect_x(np.transpose(occ, (0, 2, 1)))), (0, 2, 1))).max() \
        <= 1.5 / 65535
    assert np.abs(wm[..., 2] - np.transpose(_q16(_expect_x(np.transpose(occ, (2, 1, 0)))), (2, 1, 0))).max() \
        <= 1.5 / 65535
    # monotone along every axis, inside (0, 1)
    assert np.all(np.diff(wm[..., 0], axis=2) > 0) and np.all(np.diff(wm[..., 2], axis=0) > 0)
    assert wm[..., :3].min() > 0 and wm[..., :3].max() < 1


def test_weights_sum_to_the_row_length():
    for n in range(33):
        lo, hi = _weights(n, fp16=False)
        assert abs(lo * (32 - n) + hi * n - 32) < 1e-4
        assert lo >= 0.5 - 1e-7
    assert _weights(8, fp16=False)[1] == 2.0          # h = the high resolution while l >= low
    assert _weights(20, fp16=False)[0] == 0.5         # else l = low and h absorbs the rest


def test_fp16_weights_texture_changes_the_map(oracle):
    occ = np.zeros((32, 32, 32), np.uint32)
    occ[:, :, 10] = 1  # one occupied cell per x-row: l = 30/31, not exact in fp16
    a = ob.warpmap(occ, ob.warp_params(use_weights_texture=1))
    b = ob.warpmap(occ, ob.warp_params(use_weights_texture=0))
    assert np.abs(a[..., 0] - _q16(_expect_x(occ, fp16=True))).max() <= 1.5 / 65535
    assert np.abs(b[..., 0] - _q16(_expect_x(occ, fp16=False))).max() <= 1.5 / 65535
    assert not np.array_equal(a[..., 0], b[..., 0])


def test_axes_and_linear_switches(oracle):
    occ = (np.random.default_rng(9).random((32, 32, 32)) < 0.2).astype(np.uint32)
    ident = ob.warpmap(np.zeros((32, 32, 32), np.uint32))
    lin = ob.warpmap(occ, ob.warp_params(linear=1))
    assert np.array_equal(lin[..., :3], ident[..., :3])
    part = ob.warpmap(occ, ob.warp_params(axes=(1, 0, 1)))
    full = ob.warpmap(occ)
    assert np.array_equal(part[..., 1], ident[..., 1])
    assert np.array_equal(part[..., 0], full[..., 0]) and np.array_equal(part[..., 2], full[..., 2])


def test_sampling_is_linear_and_clamped(oracle):
    occ = (np.random.default_rng(4).random((32, 32, 32)) < 0.3).astype(np.uint32)
    wm = ob.warpmap(occ)
    # at a texel centre: the texel itself
    for x, y, z in [(0, 0, 0), (5, 17, 30), (31, 31, 31)]:
        tc = (np.array([x, y, z], np.float32) + 0.5) / 32
        assert np.allclose(ob.warp_sample(wm, tc), wm[z, y, x, :3], atol=1e-7)
    # halfway between two x-centres: the mean
    tc = np.array([6.0 / 32, 17.5 / 32, 30.5 / 32], np.float32)
    assert np.allclose(ob.warp_sample(wm, tc), 0.5 * (wm[30, 17, 5, :3] + wm[30, 17, 6, :3]), atol=1e-6)
    # outside the centres: clamped to the edge texels
    assert np.allclose(ob.warp_sample(wm, [0.0, 0.0, 0.0]), wm[0, 0, 0, :3], atol=1e-7)
    assert np.allclose(ob.warp_sample(wm, [1.5, -0.2, 1.0]), wm[31, 0, 31, :3], atol=1e-7)

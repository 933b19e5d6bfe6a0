"""The oracle's tessellation voxelizer (or_tess_levels / or_tess_point /
or_voxelize_tess; simpleTesselated.vert, testTesselation.tesc/.tese), pinned
by what the GL 4.5 core specification (11.2.2.1, triangles, equal_spacing,
point mode) and the .tesc formulas fix independently of the restatement:

  count      3 + sum(OL_e - 1) + the inner rings: 3 (n/2 - 1)(n/2) + 1 for
             even n, 3 ((n - 1)/2)^2 for odd n; all-1 levels: the 3 corners
  outer      OL_e - 1 equally spaced points inside edge e (u, v, w = 0)
  rings      ring k's corners project onto the two adjacent domain edges at
             distance k/n from the corner (the perpendicular construction, in
             an equilateral domain); its edges are cut into n - 2k equal parts
  levels     a right triangle with legs a, b in cubic voxels h: inner
             ceil(max(a, b)/h), outer ceil(len/h) in the .tesc's edge order,
             clamped to 64; one-voxel patches discarded
  voxelize   a triangle in one voxel layer fills only voxels it touches, with
             its albedo and normal * 0.5 + 0.5
"""
import math

import numpy as np
import pytest


def _count(lv):
    n, m = lv[0], lv[1:]
    if n == 1 and all(x == 1 for x in m):
        return 3
    c = 3 + sum(x - 1 for x in m)
    c += 3 * (n // 2 - 1) * (n // 2) + 1 if n % 2 == 0 else 3 * ((n - 1) // 2) ** 2
    return c


@pytest.mark.parametrize("lv", [[1, 1, 1, 1], [2, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3], [4, 2, 3, 5], [5, 1, 1, 1],
                                [6, 6, 1, 9], [7, 64, 2, 3], [64, 64, 64, 64], [63, 17, 40, 1]])
def test_point_counts_and_domain(oracle, lv):
    pts = oracle.tess_points(lv).astype(np.float64)
    assert len(pts) == _count(lv)
    assert np.all(pts >= -1e-7) and np.all(pts <= 1 + 1e-7)
    assert np.abs(pts.sum(1) - 1).max() < 1e-6
    assert len(np.unique(np.round(pts, 6), axis=0)) == len(pts)  # each distinct vertex once (point mode)


@pytest.mark.parametrize("lv", [[4, 2, 3, 5], [9, 7, 1, 12]])
def test_outer_edges_equally_spaced(oracle, lv):
    pts = oracle.tess_points(lv).astype(np.float64)
    interior = pts[3:3 + sum(m - 1 for m in lv[1:])]
    start = 0
    for e in range(3):
        m = lv[1 + e]
        seg = interior[start:start + m - 1]
        start += m - 1
        assert np.all(np.abs(seg[:, e]) < 1e-7)  # edge e lies on coordinate e = 0
        other = [q for q in range(3) if q != e]
        got = np.sort(seg[:, other[0]])
        want = np.sort(np.array([s / m for s in range(1, m)]))
        assert np.allclose(got, want, atol=1e-6) or np.allclose(np.sort(1 - seg[:, other[0]]), want, atol=1e-6)


def _to_plane(uvw):
    A, B, C = np.array([0.0, 0.0]), np.array([1.0, 0.0]), np.array([0.5, math.sqrt(3) / 2])
    return uvw[..., 0:1] * A + uvw[..., 1:2] * B + uvw[..., 2:3] * C


@pytest.mark.parametrize("n", [2, 3, 4, 7, 10, 64])
def test_inner_rings_follow_the_perpendicular_construction(oracle, n):
    lv = [n, 3, 3, 3]
    pts = oracle.tess_points(lv).astype(np.float64)
    j = 3 + 3 * 2
    corners = np.eye(3)
    for k in range(1, n // 2 + 1):
        m = n - 2 * k
        ring = pts[j:j + (1 if m == 0 else 3 * m)]
        j += len(ring)
        if m == 0:
            assert np.allclose(ring[0], [1 / 3] * 3, atol=1e-6)
            continue
        xy = _to_plane(ring)
        for e in range(3):  # ring corner near domain corner e: projections onto both adjacent edges at k/n
            rc = xy[e * m]
            c0 = _to_plane(corners[e])
            for f in range(3):
                if f == e:
                    continue
                d = _to_plane(corners[f]) - c0
                assert abs(np.dot(rc - c0, d) - k / n) < 1e-5, (n, k, e, f)
            # the ring edge from this corner to the next: m equal steps
            nxt = xy[((e + 1) % 3) * m]
            steps = xy[e * m:e * m + m] if e < 2 else np.vstack([xy[2 * m:3 * m]])
            for s in range(m):
                assert np.allclose(steps[s], rc + (nxt - rc) * (s / m), atol=1e-5)


def test_levels_of_a_right_triangle(oracle):
    a, b = 0.3, 0.2
    h = 2.0 / 64
    P = [[0.05, 0.05, 0.01], [0.05 + a, 0.05, 0.01], [0.05, 0.05 + b, 0.01]]
    n, lv = oracle.tess_levels(P, 64, [0, 0, 0], [-1, -1, -1], [1, 1, 1])
    hyp = math.hypot(a, b)
    # inner: the largest altitude (a leg); OL0: b - a (leg a), OL1: c - b (hypotenuse), OL2: c - a (leg b)
    assert lv == [math.ceil(max(a, b) / h), math.ceil(a / h), math.ceil(hyp / h), math.ceil(b / h)]
    assert n == _count(lv)
    big = [[-0.9, -0.9, 0.0], [0.9, -0.9, 0.0], [-0.9, 0.9, 0.0]]
    n, lv = oracle.tess_levels(big, 256, [0, 0, 0], [-1, -1, -1], [1, 1, 1])
    assert lv == [64, 64, 64, 64]
    tiny = [[0.101, 0.101, 0.101], [0.102, 0.101, 0.101], [0.101, 0.102, 0.101]]
    assert oracle.tess_levels(tiny, 64, [0, 0, 0], [-1, -1, -1], [1, 1, 1])[0] == 0


def test_voxelized_triangle_fills_only_its_layer(oracle):
    dim = 32
    z = 0.0 + 0.5 * (2.0 / dim)  # the centre of a voxel layer
    pos = np.array([[[-0.5, -0.5, z], [0.5, -0.5, z], [-0.5, 0.5, z]]], np.float32)
    nrm = np.array([[[0, 0, 1]] * 3], np.float32)
    albedo = np.array([0xff336699], np.uint32)
    mesh = dict(pos=pos, nrm=nrm, tan=np.array([[1, 0, 0]], np.float32), albedo=albedo,
                obj=np.zeros(1, np.uint32), rough=None)
    sc = oracle.SceneArrays(mesh, np.eye(4, dtype=np.float32).T.reshape(1, 16))
    alb, nr = oracle.voxelize_tess(sc, dim, [0, 0, 0], [-1, -1, -1], [1, 1, 1], accum=1)
    occ = np.argwhere(alb != 0)
    layer = int(dim * ((z + 1) / 2))
    assert len(occ) > 0 and np.all(occ[:, 0] == layer)
    assert np.all(alb[alb != 0] == albedo[0])
    want_n = 0 | (int(round(0.5 * 255)) << 0) | (int(round(0.5 * 255)) << 8) | (255 << 16) | (255 << 24)
    assert np.all(nr[alb != 0] == want_n)
    # every filled voxel touches the triangle x + y <= 0 (with its voxel's slack), inside [-0.5, 0.5]^2
    cx = (occ[:, 2] + 0.5) * (2.0 / dim) - 1
    cy = (occ[:, 1] + 0.5) * (2.0 / dim) - 1
    hv = 1.0 / dim
    assert np.all(cx >= -0.5 - hv) and np.all(cy >= -0.5 - hv) and np.all(cx + cy <= 0.0 + 2 * hv)

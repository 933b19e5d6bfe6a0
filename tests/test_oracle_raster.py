"""Pins the oracle's rasterization rules against an independent exact
restatement.

The reference rasterizes through GL (shaders/voxelize.geom -> fixed-function
raster -> voxelize.frag, src/Application.cpp:666-754).  The oracle restates
GL's rules with integer edge functions; this test restates them again with
exact rational geometry instead: a sample is covered when it is strictly
inside the (snapped) triangle, or lies on a top edge (horizontal, third vertex
below) or a left edge (third vertex to its right).  Triangles have vertices
on a 1/8-pixel lattice so samples fall exactly on edges and vertices often.
Coverage is observed through the voxelizer (4x MSAA, any covered sample makes
a fragment) on a constant-depth plane, where the voxel (x, y) is the pixel."""
from fractions import Fraction as Fr

import numpy as np
import pytest

import vct_amd
from vct_amd import _abi as abi

DIM = 16
SAMPLES_MSAA = [(Fr(3, 8), Fr(1, 8)), (Fr(7, 8), Fr(3, 8)), (Fr(1, 8), Fr(5, 8)), (Fr(5, 8), Fr(7, 8))]


def _side(a, b, p):
    return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0])


def covered(tri, p):
    a, b, c = tri
    area = _side(a, b, c)
    if area == 0:
        return False
    if area < 0:
        a, b = b, a
    for (u, v, w) in ((a, b, c), (b, c, a), (c, a, b)):
        s = _side(u, v, p)
        if s < 0:
            return False
        if s == 0:
            if u[1] == v[1]:
                if not w[1] < u[1]:  # a top edge has the third vertex below it
                    return False
            else:
                xl = u[0] + (w[1] - u[1]) * (v[0] - u[0]) / (v[1] - u[1])
                if not w[0] > xl:  # a left edge has the third vertex to its right
                    return False
    return True


def expected_pixels(tri, msaa=True):
    pix = set()
    xs = [v[0] for v in tri]
    ys = [v[1] for v in tri]
    for py in range(max(0, int(min(ys)) - 1), min(DIM, int(max(ys)) + 2)):
        for px in range(max(0, int(min(xs)) - 1), min(DIM, int(max(xs)) + 2)):
            samples = SAMPLES_MSAA if msaa else [(Fr(1, 2), Fr(1, 2))]
            if any(covered(tri, (px + sx, py + sy)) for sx, sy in samples):
                pix.add((px, py))
    return pix


def run_oracle(oracle, tris_win, msaa):
    """tris_win: window-space triangles (pixel units, exact on the 1/256 grid).
    They are placed in the plane z = depth of voxel layer 5 and voxelized."""
    tris = np.array(tris_win, np.float64)
    # window -> world for the z-axis projection of the [-1,1]^3 grid: x = 2*xw/DIM - 1
    pos = np.zeros((len(tris), 3, 3), np.float32)
    pos[:, :, 0] = (2.0 * tris[:, :, 0] / DIM - 1.0)
    pos[:, :, 1] = (2.0 * tris[:, :, 1] / DIM - 1.0)
    pos[:, :, 2] = -1.0 + (5.5 * 2.0 / DIM)
    nrm = np.zeros_like(pos)
    nrm[:, :, 2] = 1.0
    mesh = dict(pos=pos, nrm=nrm, tan=np.tile([1, 0, 0], (len(tris), 1)).astype(np.float32),
                albedo=np.full(len(tris), 0xff808080, np.uint32), obj=np.zeros(len(tris), np.uint32), nobj=1)
    model = np.eye(4, dtype=np.float32).reshape(1, 16)
    sc = oracle.SceneArrays(mesh, model)
    vox = abi.VoxelizeParams()
    mx, my, mz = oracle.voxel_projections((0, 0, 0), (-1, -1, -1), (1, 1, 1))
    vox.mvp_x[:], vox.mvp_y[:], vox.mvp_z[:] = list(mx), list(my), list(mz)
    vox.axis_override, vox.lighting, vox.atomic_max, vox.msaa = 2, 0, 1, msaa
    alb, _, st = oracle.voxelize(sc, [], [], vox, DIM)
    occ = (alb >> 24) > 0
    assert occ.sum() == occ[5].sum()  # everything landed in the depth layer of the plane
    return {(int(x), int(y)) for y, x in zip(*np.nonzero(occ[5]))}, st


@pytest.mark.parametrize("msaa", [1, 0])
def test_coverage_matches_exact_rational_raster(oracle, msaa):
    rng = np.random.default_rng(7 + msaa)
    for trial in range(150):
        # vertices on the 1/8 pixel lattice -> frequent exact ties at samples
        v = rng.integers(0, DIM * 8 + 1, size=(3, 2))
        tri = [(Fr(int(a), 8), Fr(int(b), 8)) for a, b in v]
        got, _ = run_oracle(oracle, [[(float(x), float(y)) for x, y in tri]], msaa)
        assert got == expected_pixels(tri, msaa=bool(msaa)), (trial, tri)


def test_shared_edges_cover_each_sample_once(oracle):
    """A square split along a diagonal: fragments = sum over both triangles of
    pixels with a covered sample; with the top-left rule no sample is counted
    twice, so a pixel is in both triangles only when they split its samples."""
    sq = [(Fr(2), Fr(2)), (Fr(12), Fr(3)), (Fr(11), Fr(13)), (Fr(3), Fr(11))]
    t1, t2 = [sq[0], sq[1], sq[2]], [sq[0], sq[2], sq[3]]
    _, st = run_oracle(oracle, [[(float(x), float(y)) for x, y in t] for t in (t1, t2)], 1)
    exp = len(expected_pixels(t1)) + len(expected_pixels(t2))
    assert st.total_voxel_fragments == exp
    for py in range(DIM):
        for px in range(DIM):
            for sx, sy in SAMPLES_MSAA:
                p = (px + sx, py + sy)
                assert not (covered(t1, p) and covered(t2, p))


def test_shuffle_invariance_and_popcount(oracle):
    """atomicMax on packed words makes voxelization order-independent
    (voxelize.frag:271-273) and uniqueVoxels is the occupancy popcount."""
    app = vct_amd.App(1, device=-1)
    mesh, fs = app.mesh(), app.frame_state(0)
    sc = oracle.SceneArrays(mesh, fs["model"])
    L = fs["lights"]
    sm = oracle.shadowmap(sc, L[0])
    a1, n1, s1 = oracle.voxelize(sc, L, [sm], fs["vox"], 64)
    perm = np.random.default_rng(3).permutation(len(mesh["albedo"]))
    mesh2 = {k: (v[perm] if isinstance(v, np.ndarray) and len(v) == len(perm) else v) for k, v in mesh.items()}
    sc2 = oracle.SceneArrays(mesh2, fs["model"])
    sm2 = oracle.shadowmap(sc2, L[0])
    assert np.array_equal(sm, sm2)
    a2, n2, s2 = oracle.voxelize(sc2, L, [sm2], fs["vox"], 64)
    assert np.array_equal(a1, a2) and np.array_equal(n1, n2)
    assert s1.total_voxel_fragments == s2.total_voxel_fragments
    e = 1 + 1 / 63
    _, _, _, st = oracle.normalize_inject(a1, 64, (0, 0, 0), (-e,) * 3, (e,) * 3, L, [sm])
    assert st.unique_voxels == int(((a1 >> 24) > 0).sum())
    assert st.max_fragments_per_voxel == 255  # alpha byte of packUnorm4x8(vec4(c, 1))


def test_axis_aligned_walls_land_on_their_voxel_layer(oracle):
    """Scene 1's room walls at +-1 sit on voxel centres of the +-64/63 extent:
    each wall occupies exactly one full 62x62 (interior) layer of voxels."""
    app = vct_amd.App(1, device=-1)
    mesh, fs = app.mesh(), app.frame_state(0)
    room = {k: (v[:12] if isinstance(v, np.ndarray) and len(v) == len(mesh["albedo"]) else v)
            for k, v in mesh.items()}
    sc = oracle.SceneArrays(room, fs["model"])
    alb, _, _ = oracle.voxelize(sc, fs["lights"], [None], fs["vox"], 64, lighting=0)
    occ = (alb >> 24) > 0  # [z, y, x]
    for layer in (occ[0], occ[63], occ[:, 0], occ[:, 63], occ[:, :, 0], occ[:, :, 63]):
        assert layer.all()
    assert occ[1:63, 1:63, 1:63].sum() == 0

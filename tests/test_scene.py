"""The seeded synthetic scenes of the application layer (include/vct_app.h)."""
import numpy as np
import pytest

import vct_amd


@pytest.fixture(scope="module")
def s1():
    return vct_amd.App(1, device=-1)


@pytest.fixture(scope="module")
def s2():
    return vct_amd.App(2, device=-1)


def _winding_normals(pos):
    e1 = pos[:, 1] - pos[:, 0]
    e2 = pos[:, 2] - pos[:, 0]
    return np.cross(e1.astype(np.float64), e2.astype(np.float64))


def test_scene1_counts_and_orientation(s1):
    nt, no, nl = s1.sizes()
    assert (nt, no, nl) == (36, 3, 1)
    m = s1.mesh()
    assert m["rough"] is None
    cr = _winding_normals(m["pos"])
    n = m["nrm"][:, 0].astype(np.float64)
    # counter-clockwise seen from the side the normal points to
    assert np.all((cr * n).sum(1) > 0)
    # the room (object 0) faces inward, the boxes outward
    centroid = m["pos"].mean(1)
    room = m["obj"] == 0
    assert room.sum() == 12
    assert np.all((n[room] * centroid[room]).sum(1) < 0)
    assert np.all((n[~room] * centroid[~room]).sum(1) > 0)  # object space: boxes centred on the origin
    t = m["tan"].astype(np.float64)
    assert np.allclose(np.linalg.norm(t, axis=1), 1, atol=1e-6)
    assert np.allclose((t * n).sum(1), 0, atol=1e-6)
    a = m["albedo"]
    assert np.all(a >> 24 == 255)
    for k in range(3):
        ch = (a >> (8 * k)) & 255
        assert ch.min() >= 51 and ch.max() <= 230  # U[0.2, 0.9] * 255, rounded


def test_scene1_deterministic(s1):
    other = vct_amd.App(1, device=-1)
    a, b = s1.mesh(), other.mesh()
    for k in ("pos", "nrm", "tan", "albedo", "obj"):
        assert np.array_equal(a[k], b[k])
    fa, fb = s1.frame_state(0), other.frame_state(0)
    assert np.array_equal(fa["model"], fb["model"])
    third = vct_amd.App(1, device=-1, seed=99)
    assert not np.array_equal(third.mesh()["pos"], a["pos"])


def test_scene1_light_and_camera(s1):
    fs = s1.frame_state(0)
    L = fs["lights"][0]
    d = np.array(list(L.direction))
    assert abs(np.linalg.norm(d) - 1) < 1e-6 and d[1] < 0  # from the upper hemisphere
    assert L.shadow_caster == 1 and L.shadow_size == 512 and L.type == 1
    ls = np.array(list(L.ls)).reshape(4, 4).T
    inv = np.array(list(L.ls_inverse)).reshape(4, 4).T
    assert np.allclose(ls @ inv, np.eye(4), atol=1e-5)
    # the fitted light frustum contains the room
    for c in np.array(np.meshgrid([-1, 1], [-1, 1], [-1, 1])).T.reshape(-1, 3):
        p = ls @ np.r_[c, 1]
        assert np.all(np.abs(p[:3]) <= 1)
    cam = fs["cam"]
    assert (cam.width, cam.height) == (256, 256)
    assert np.all(np.abs(np.array(list(cam.eye))) < 0.6)


def test_scene2_counts(s2):
    nt, no, nl = s2.sizes()
    assert nt == 2000 * 12 + 64 * 20 * 4 ** 4 == 351_680
    assert no == 2064 and nl == 5
    m = s2.mesh()
    assert m["rough"] is not None and m["rough"].min() >= 0.1 and m["rough"].max() <= 0.8
    cr = _winding_normals(m["pos"])
    n = m["nrm"].astype(np.float64).mean(1)
    assert np.all((cr * n).sum(1) > 0)
    sph = m["obj"] >= 2000
    assert sph.sum() == 64 * 5120
    r = np.linalg.norm(m["pos"][sph].astype(np.float64), axis=2)
    rad = np.array([r[m["obj"][sph] == o].mean() for o in range(2000, 2064)])
    assert rad.min() >= 0.02 - 1e-6 and rad.max() <= 0.1 + 1e-6
    fs = s2.frame_state(0)
    types = [L.type for L in fs["lights"]]
    assert types == [1, 0, 0, 0, 0]
    assert all(L.shadow_size == 1024 for L in fs["lights"])
    assert (fs["cam"].width, fs["cam"].height) == (1920, 1080)
    assert fs["trace"].enable_reflections == 1


def test_scene2_motion_is_rigid_and_bounded(s2):
    f0, f1, f50 = (s2.frame_state(f)["model"] for f in (0, 1, 50))
    assert not np.array_equal(f0, f1)
    for m in (f0, f50):
        M = m.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
        R = M[:, :3, :3]
        assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
        assert np.allclose(np.linalg.det(R), 1, atol=1e-5)
        assert np.all(np.abs(M[:, :3, 3]) <= 0.95 + 1e-6)
    # per-frame motion: rotation <= 5 degrees, translation <= 0.01 per axis
    M0 = f0.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    M1 = f1.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    rel = M1[:, :3, :3] @ M0[:, :3, :3].transpose(0, 2, 1)
    ang = np.degrees(np.arccos(np.clip((np.trace(rel, axis1=1, axis2=2) - 1) / 2, -1, 1)))
    assert ang.max() <= 5.0 + 1e-3
    assert np.abs(M1[:, :3, 3] - M0[:, :3, 3]).max() <= 0.01 + 1e-6


@pytest.fixture(scope="module")
def s5():
    return vct_amd.App(5, device=-1)


def test_scene5_counts_and_config(s5):
    """Scene 5: scene 2's distributions at 4x the extent -- 8000 boxes + 512
    subdivision-4 icospheres (2,717,440 triangles) in [-4,4]^3 with sizes x2.5,
    8 lights, a 3840x2160 camera, FP8 coarse levels and the warp texture at
    4.0."""
    nt, no, nl = s5.sizes()
    assert nt == 8000 * 12 + 512 * 20 * 4 ** 4 == 2_717_440
    assert no == 8512 and nl == 8
    c = s5.cfg
    assert (c.dim, c.levels, c.anisotropic, c.storage) == (256, 8, 1, 2)
    assert (c.width, c.height, c.reflections) == (3840, 2160, 1)
    assert c.warp_texture == 1 and c.warp_high_resolution == 4.0
    m = s5.mesh()
    sph = m["obj"] >= 8000
    assert sph.sum() == 512 * 5120
    r = np.linalg.norm(m["pos"][sph].astype(np.float64), axis=2)
    rad = np.array([r[m["obj"][sph] == o].mean() for o in range(8000, 8512, 16)])
    assert rad.min() >= 0.05 - 1e-5 and rad.max() <= 0.25 + 1e-5
    box = ~sph
    half = np.abs(m["pos"][box].astype(np.float64)).max(axis=1)  # per-triangle max |coordinate|
    # half-extents 2.5 lognormal(-3, 0.5): median 2.5 e^-3 ~ 0.12
    assert 0.06 < np.median(half) < 0.25
    # the boxes fill about the same fraction of the extent as in scene 2: the
    # mean box volume is 8 Z^3 E[e^X]^3 with X ~ N(-3, 0.5^2)
    frac5 = 8 * 2.5 ** 3 * np.exp(3 * (-3.0 + 0.25 / 2)) * 8000 / 8 ** 3
    frac2 = 8 * np.exp(3 * (-3.0 + 0.25 / 2)) * 2000 / 2 ** 3
    assert abs(frac5 / frac2 - 1) < 0.05 and 0.2 < frac5 < 0.5
    fs = s5.frame_state(0)
    assert [L.type for L in fs["lights"]] == [1] + [0] * 7
    assert all(L.shadow_caster and L.shadow_size == 1024 for L in fs["lights"])
    assert all(np.abs(np.asarray(L.position[:])).max() <= 3.2 + 1e-5 for L in fs["lights"][1:])
    assert (fs["cam"].width, fs["cam"].height) == (3840, 2160)
    assert fs["vox"].warp_texture == 1 and fs["inj"].warp_texture == 1 and fs["trace"].warp_texture == 1


def test_scene5_motion_is_bounded_by_the_extent(s5):
    f0, f1, f80 = (s5.frame_state(f)["model"] for f in (0, 1, 80))
    for m in (f0, f80):
        M = m.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
        assert np.all(np.abs(M[:, :3, 3]) <= 3.8 + 1e-5)
    M0 = f0.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    M1 = f1.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    assert np.abs(M1[:, :3, 3] - M0[:, :3, 3]).max() <= 0.04 + 1e-5

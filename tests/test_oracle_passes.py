"""Pins the oracle's per-pass restatements by brute force and closed forms.

  inject       brute-force per-texel unprojection in numpy (injectRadiance.comp:
               152-168) + per-voxel closed form (transferVoxels.comp:37-62)
  fill holes   brute-force 3x3x3 neighbourhood mean (voxelFillHoles.comp)
  mip          constant and ramp closed forms, the BOX2 offset order in numpy
               float32, fp16 rounding against numpy
  aniso mip    uniform-volume closed form of the front-to-back composite
  trace        empty volume -> (0, occlusion 1); uniform volume closed form
               alpha_n = 1 - (1 - a)^n, c_n = c (1 - (1 - a)^n) / a with n from
               the h_{k+1} = h_k (1 + tan(theta/2)) schedule; FP32 vs FP64
  G-buffer     one quad facing the camera: positions are ray-plane hits
  temporal     decay 0 = the plain pass bit for bit; decay 1 keeps every unlit
               voxel; lit voxels always equal the plain pass; on a static scene
               the alpha of an unlit occupied voxel follows a (1 - decay^k)
  camera snap  voxelTrackCamera: centre = floor(eye / cell) * cell with
               cell = 2^levels (max - min) / dim, seen through the voxel
               projections the application derives from it
"""
import math

import numpy as np
import pytest

import vct_amd
from vct_amd import _abi as abi


@pytest.fixture(scope="module")
def scene1(oracle):
    app = vct_amd.App(1, device=-1)
    mesh, fs = app.mesh(), app.frame_state(0)
    sc = oracle.SceneArrays(mesh, fs["model"])
    L = fs["lights"]
    sm = oracle.shadowmap(sc, L[0])
    alb, _, _ = oracle.voxelize(sc, L, [sm], fs["vox"], 64)
    e = np.float32(1 + 1 / 63)
    return dict(app=app, mesh=mesh, fs=fs, sc=sc, L=L, sm=sm, alb=alb, ext=((0, 0, 0), (-e,) * 3, (e,) * 3))


def _f32(x):
    return np.float32(x)


def test_inject_brute_force(oracle, scene1):
    L, sm, alb = scene1["L"][0], scene1["sm"], scene1["alb"]
    c, mn, mx = (np.asarray(v, np.float32) for v in scene1["ext"])
    alb2, rad0, lm, st = oracle.normalize_inject(alb, 64, *scene1["ext"], scene1["L"], [sm])
    N = sm.shape[0]
    inv = np.asarray(list(L.ls_inverse), np.float32).reshape(4, 4).T
    expect = np.zeros(64 ** 3, np.uint8)
    smp = np.pad(sm, 1, constant_values=1.0)  # border depth 1.0
    for j in range(N):
        for i in range(N):
            # tc = (i, j) / N is a texel corner: GL LINEAR averages the 2x2 texels around it
            d = _f32(0.25) * (smp[j, i] + smp[j, i + 1] + smp[j + 1, i] + smp[j + 1, i + 1])
            ndc = np.array([_f32(i) / _f32(N) * 2 - 1, _f32(j) / _f32(N) * 2 - 1, d * 2 - 1, 1], np.float32)
            w = (inv.astype(np.float64) @ ndc.astype(np.float64))[:3]
            idx = np.trunc(64 * (w - c - mn) / (mx - mn)).astype(int)
            if (idx >= 0).all() and (idx < 64).all():
                expect[(idx[2] * 64 + idx[1]) * 64 + idx[0]] = 1
    diff = np.flatnonzero(expect != lm)
    # float64 unprojection vs the oracle's float32: only voxel-boundary ties may differ
    assert len(diff) <= max(3, int(0.002 * expect.sum())), len(diff)
    occ = (alb.reshape(-1) >> 24) > 0
    lit = lm.astype(bool)
    r = rad0.reshape(-1, 4)
    assert np.all(r[~occ] == 0)
    assert np.all(r[occ, 3] == 0.5)
    assert np.all(r[occ & ~lit, :3] == 0)
    rgb = np.stack([(alb.reshape(-1) >> (8 * k)) & 255 for k in range(3)], 1)[occ & lit]
    ref = (rgb.astype(np.float32) / np.float32(255)).astype(np.float16).astype(np.float32)
    assert np.array_equal(r[occ & lit, :3], ref)
    assert np.all((alb2.reshape(-1)[occ] >> 24) == 128)  # voxelSetOpacity 0.5 -> 128/255
    assert st.lit_voxels == int((occ & lit).sum())


def test_inject_is_idempotent(oracle, scene1):
    s = scene1
    a1, r1, l1, _ = oracle.normalize_inject(s["alb"], 64, *s["ext"], s["L"], [s["sm"]])
    a2, r2, l2, _ = oracle.normalize_inject(a1, 64, *s["ext"], s["L"], [s["sm"]])
    assert np.array_equal(r1, r2) and np.array_equal(l1, l2) and np.array_equal(a1, a2)


def _inject(oracle, s, prev=None, decay=0.8):
    return oracle.normalize_inject(s["alb"], 64, *s["ext"], s["L"], [s["sm"]], prev=prev, decay=decay)


def test_temporal_decay_zero_is_the_plain_pass(oracle, scene1):
    rng = np.random.default_rng(5)
    prev = rng.uniform(0, 1, (64 ** 3, 4)).astype(np.float16).astype(np.float32)
    a0, r0, l0, s0 = _inject(oracle, scene1)
    a1, r1, l1, s1 = _inject(oracle, scene1, prev=prev, decay=0.0)
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
    assert np.array_equal(a0, a1) and np.array_equal(l0, l1)
    assert (s0.unique_voxels, s0.lit_voxels) == (s1.unique_voxels, s1.lit_voxels)


def test_temporal_decay_one_keeps_unlit_voxels(oracle, scene1):
    rng = np.random.default_rng(6)
    prev = rng.uniform(0, 1, (64 ** 3, 4)).astype(np.float16).astype(np.float32)
    _, plain, lm, _ = _inject(oracle, scene1)
    _, r, lm2, _ = _inject(oracle, scene1, prev=prev, decay=1.0)
    lit = lm.astype(bool)
    r, plain = r.reshape(-1, 4), plain.reshape(-1, 4)
    assert np.array_equal(r[~lit], prev[~lit])      # mix(prev, c, 0) = prev
    assert np.array_equal(r[lit], plain[lit])       # lit voxels are overwritten
    assert lit.sum() > 100


def test_temporal_static_scene_converges_geometrically(oracle, scene1):
    d = 0.8
    _, plain, lm, _ = _inject(oracle, scene1)
    plain = plain.reshape(-1, 4)
    occ = (scene1["alb"].reshape(-1) >> 24) > 0
    lit = lm.astype(bool)
    prev = np.zeros((64 ** 3, 4), np.float32)
    for k in range(1, 9):
        _, r, _, _ = _inject(oracle, scene1, prev=prev, decay=d)
        r = r.reshape(-1, 4)
        a = r[occ & ~lit, 3].astype(np.float64)
        # alpha_k = 0.5 (1 - d^k), each frame rounded to fp16 (relative 2^-11 per step)
        assert np.allclose(a, 0.5 * (1 - d ** k), rtol=k * 2 ** -10, atol=0)
        assert np.all(r[occ & ~lit, :3] == 0)       # never-lit voxels gain no colour
        assert np.all(r[~occ & ~lit] == 0)
        assert np.array_equal(r[lit], plain[lit])
        prev = r


def test_track_camera_snaps_the_grid_centre(oracle):
    import vct_amd
    a0 = vct_amd.App(1, device=-1)
    a1 = vct_amd.App(1, device=-1, track_camera=1)
    fs = a1.frame_state(0)
    eye = np.asarray(list(fs["cam"].eye), np.float32)
    e = np.float32(1 + 1 / 63)
    cell = np.float32(2.0 ** 7) * (e - (-e)) / np.float32(64)
    centre = np.floor(eye / cell) * cell
    expect = oracle.voxel_projections(tuple(centre), (-e,) * 3, (e,) * 3)
    got = [np.asarray(list(getattr(fs["vox"], m)), np.float32) for m in ("mvp_x", "mvp_y", "mvp_z")]
    for g, x in zip(got, expect):
        assert np.array_equal(g.reshape(-1), np.asarray(x, np.float32).reshape(-1))
    base = [np.asarray(list(getattr(a0.frame_state(0)["vox"], m)), np.float32) for m in ("mvp_x",)]
    assert np.any(centre != 0) == (not np.array_equal(base[0], got[0]))


def test_fill_holes_brute_force(oracle):
    rng = np.random.default_rng(0)
    d = 8
    v = np.zeros((d, d, d, 4), np.float32)
    occ = rng.random((d, d, d)) < 0.3
    v[occ] = rng.random((occ.sum(), 4)).astype(np.float32)
    v[occ, 3] = 0.5
    got = oracle.fill_holes(v, fp32=True)
    for z in range(d):
        for y in range(d):
            for x in range(d):
                if v[z, y, x, 3] > 0:
                    assert np.array_equal(got[z, y, x], v[z, y, x])
                    continue
                nb = v[max(0, z - 1):z + 2, max(0, y - 1):y + 2, max(0, x - 1):x + 2].reshape(-1, 4)
                nb = nb[nb[:, 3] > 0]
                exp = nb.astype(np.float64).mean(0) if len(nb) else np.zeros(4)
                assert np.allclose(got[z, y, x], exp, atol=1e-6)


def test_mip_closed_forms(oracle):
    d, L = 32, 6
    const = np.full((d, d, d, 4), [0.25, 0.5, 0.75, 0.5], np.float32)
    lv = oracle.mip(const, L)
    for l in range(1, L):
        assert lv[l].shape[0] == d >> l
        assert np.all(lv[l] == const[0, 0, 0])
    # linear ramp in x: the 2x2x2 mean is the value at the block centre
    x = np.arange(d, dtype=np.float32) / 64
    ramp = np.zeros((d, d, d, 4), np.float32)
    ramp[..., 0] = x[None, None, :]
    lv = oracle.mip(ramp, 3, fp32=True)
    xs = (np.arange(d // 2) * 2 + 0.5) / 64
    assert np.allclose(lv[1][0, 0, :, 0], xs, atol=1e-7)


def test_mip_box2_order_and_fp16_rounding(oracle):
    rng = np.random.default_rng(1)
    d = 16
    v = rng.random((d, d, d, 4)).astype(np.float16).astype(np.float32)
    lv = oracle.mip(v, 4)
    src = v
    for l in range(1, 4):
        s = src.shape[0] // 2
        acc = np.zeros((s, s, s, 4), np.float32)
        # filterRadiance.comp BOX2 offsets: (0,0,0),(0,0,1),(0,1,0),(0,1,1),(1,0,0),... as (x,y,z)
        for ox, oy, oz in [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]:
            acc = acc + src[oz::2, oy::2, ox::2]
        ref = (acc * np.float32(0.125)).astype(np.float16).astype(np.float32)
        assert np.array_equal(lv[l], ref), l
        src = ref


def test_mip_mean_is_preserved_in_fp32(oracle):
    rng = np.random.default_rng(2)
    v = rng.random((32, 32, 32, 4)).astype(np.float32)
    lv = oracle.mip(v, 6, fp32=True)
    for l in range(1, 6):
        assert abs(lv[l].astype(np.float64).mean() - v.astype(np.float64).mean()) < 1e-6


def test_aniso_uniform_closed_form(oracle):
    d, L = 16, 5
    c, a = 0.3, 0.4
    v = np.zeros((d, d, d, 4), np.float32)
    v[..., :3] = c
    v[..., 3] = a
    faces = oracle.mip_aniso(v, L, fp32=True)
    cc, aa = c, a
    for l in range(1, L):
        cc, aa = cc + (1 - aa) * cc, aa + (1 - aa) * aa  # near + (1 - near.a) * far, near == far
        for f in range(6):
            assert np.allclose(faces[l][f][..., 0], cc, atol=1e-6)
            assert np.allclose(faces[l][f][..., 3], aa, atol=1e-6)


def test_aniso_faces_see_occluders_in_order(oracle):
    """An opaque red slab in front of a green one along +x: the +x face (rays
    travelling +x meet red first) is red, the -x face is green."""
    v = np.zeros((2, 2, 2, 4), np.float32)
    v[:, :, 0] = [1, 0, 0, 1]
    v[:, :, 1] = [0, 1, 0, 1]
    f = oracle.mip_aniso(v, 2, fp32=True)
    assert np.allclose(f[1][0][0, 0, 0], [1, 0, 0, 1])
    assert np.allclose(f[1][1][0, 0, 0], [0, 1, 0, 1])


def _one_level_trace(oracle, vol, dirn, pos=(0.5, 0.5, 0.5), nrm=(0, 0, 1), steps=16, angle=math.radians(60)):
    import ctypes as C
    lp = (C.c_void_p * 1)(vol.ctypes.data)
    out = np.zeros(4, np.float32)
    p, n, dd = (np.asarray(t, np.float32) for t in (pos, nrm, dirn))
    cone = oracle.Cone(steps, angle, 1.0, 1.0, 0.5)
    oracle.lib().or_trace_cone(C.cast(lp, C.c_void_p), None, vol.shape[0], 1, p.ctypes.data, n.ctypes.data,
                               dd.ctypes.data, C.byref(cone), out.ctypes.data)
    return out


def test_trace_empty_volume(oracle):
    vol = np.zeros((64, 64, 64, 4), np.float32)
    assert np.all(_one_level_trace(oracle, vol, (1, 0, 0)) == 0)


@pytest.mark.parametrize("a", [0.5, 0.2, 0.05])
def test_trace_uniform_volume_closed_form(oracle, a):
    c = 0.3
    vol = np.zeros((64, 64, 64, 4), np.float32)
    vol[..., :3] = c
    vol[..., 3] = a
    got = _one_level_trace(oracle, vol, (0, 0, 1), steps=16)
    # sample schedule: h_0 = 1, h_{k+1} = h_k (1 + tan 30deg); stop when alpha >= 0.95,
    # after `steps` samples, or when the sample leaves the volume (or reaches its
    # border texels, where trilinear filtering mixes in the zero border)
    t = math.tan(math.radians(30))
    h, n, alpha = 1.0, 0, 0.0
    start = 0.5 + 1.0 / 64
    while n < 16 and alpha < 0.95:
        z = start + h / 64
        if z > 1 - 0.5 / 64:
            break
        n += 1
        alpha = 1 - (1 - a) ** n
        h += h * t
    assert abs(got[3] - alpha) < 1e-5
    assert abs(got[0] - c * alpha / a) < 1e-5


def test_trace_fp32_matches_fp64(oracle, scene1):
    s = scene1
    alb2, rad0, _, _ = oracle.normalize_inject(s["alb"], 64, *s["ext"], s["L"], [s["sm"]])
    levels = oracle.mip(rad0, 7)
    gb = oracle.gbuffer(s["sc"], s["fs"]["cam"])
    tp = oracle.trace_params(s["fs"]["trace"], 64, 7, *s["ext"], False)
    i32, _ = oracle.trace(gb, levels, tp)
    i64, _ = oracle.trace(gb, levels, tp, fp64=True)
    err = np.abs(i32[:, :3] - i64[:, :3]).sum() / np.abs(i64[:, :3]).sum()
    assert err < 1e-5
    assert i64[:, :3].sum() > 0


def test_gbuffer_quad_positions_are_ray_plane_hits(oracle):
    W = H = 32
    z0 = -2.0
    q = np.array([[-5, -5, z0], [5, -5, z0], [5, 5, z0], [-5, 5, z0]], np.float32)
    pos = np.stack([q[[0, 1, 2]], q[[0, 2, 3]]])
    nrm = np.zeros_like(pos)
    nrm[..., 2] = 1
    mesh = dict(pos=pos, nrm=nrm, tan=np.array([[1, 0, 0]] * 2, np.float32),
                albedo=np.array([0xff0000ff, 0xff00ff00], np.uint32), obj=np.zeros(2, np.uint32), nobj=1)
    sc = oracle.SceneArrays(mesh, np.eye(4, dtype=np.float32).reshape(1, 16))
    cam = abi.Camera()
    cam.width, cam.height, cam.cull_back = W, H, 1
    cam.view[:] = list(oracle.lookat((0, 0, 0), (0, 0, -1), (0, 1, 0)))
    cam.projection[:] = list(oracle.perspective(math.radians(60), 1.0, 0.1, 10.0))
    g = oracle.gbuffer(sc, cam)
    assert np.all(g[..., 11].view(np.uint32) == 1)
    th = math.tan(math.radians(30))
    for py in range(H):
        for px in range(W):
            xn, yn = (2 * (px + 0.5) / W - 1), (2 * (py + 0.5) / H - 1)
            hit = np.array([xn * th * -z0, yn * th * -z0, z0])
            assert np.allclose(g[py, px, :3], hit, atol=2e-5)
            assert np.allclose(g[py, px, 4:7], [0, 0, 1])
            assert np.allclose(g[py, px, 8:11], [1, 0, 0], atol=1e-6)   # T
            assert np.allclose(g[py, px, 12:15], [0, 1, 0], atol=1e-6)  # B = N x T
    # culling: the same quad seen from behind is back-facing
    cam.view[:] = list(oracle.lookat((0, 0, -4), (0, 0, 0), (0, 1, 0)))
    g = oracle.gbuffer(sc, cam)
    assert np.all(g[..., 11].view(np.uint32) == 0)


def test_parallel_oracle_is_thread_count_independent(oracle, scene1):
    """The oracle's shadow map, voxelize, inject and mip loops run on OpenMP
    threads; the depth and voxel stores are atomic min/max/or, so one thread
    and many give identical bits."""
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    s = scene1
    runs = []
    for n in (1, 4):
        gomp.omp_set_num_threads(n)
        assert np.array_equal(oracle.shadowmap(s["sc"], s["L"][0]).view(np.uint32), s["sm"].view(np.uint32))
        alb, nrm, st = oracle.voxelize(s["sc"], s["L"], [s["sm"]], s["fs"]["vox"], 64)
        a2, rad0, lm, st2 = oracle.normalize_inject(alb, 64, *s["ext"], s["L"], [s["sm"]])
        lv = oracle.mip(rad0, 7)
        runs.append((alb, nrm, st.total_voxel_fragments, a2, rad0, lm, st2.unique_voxels, st2.lit_voxels, lv))
    gomp.omp_set_num_threads(max(1, __import__("os").cpu_count() or 1))
    a, b = runs
    for x, y in zip(a[:8], b[:8]):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    for x, y in zip(a[8], b[8]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))

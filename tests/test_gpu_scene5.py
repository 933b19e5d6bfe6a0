"""Scene 5, the warped 4K workload: 2,717,440 triangles in [-4,4]^3, moving
rigidly, 8 lights with shadow maps, 6-direction mips with FP8 E4M3 (+ a
per-brick scale) on levels >= 3, and the warp texture at high resolution 4.0
through which voxelize, inject and trace run (src/Application.cpp:235-575,
Overlay.cpp:289).

At reduced size (64^3, 480x270, 256^2 shadow maps) every pass is compared
with the oracle: shadow maps, warp occupancy and texture, warped voxelize and
inject bit for bit, every face of the FP8 pyramid bit for bit, the warped
trace within 1e-3 mean relative error.  At full size (256^3, 3840x2160):
determinism of the whole frame, the statistics' consistency, and the warp
texture against the oracle's warp rule applied to the GPU's occupancy."""
import numpy as np
import pytest

from gpu_util import make_app, rel_err
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def _warp_params():
    p = abi.default_warpmap_params()
    p.high_resolution = 4.0
    return p


@pytest.fixture(scope="module")
def reduced():
    app = make_app(5, dim=64, levels=7, width=480, height=270, shadow_size=256)
    ctx = app.ctx
    fs = app.frame_state(4)
    g = ctx.grid()
    assert g.storage == abi.VCT_STORE_FP8 and g.anisotropic
    print("scene 5 @64^3: shadow maps, warpmap, voxelize, inject, mip, trace", flush=True)
    ctx.set_transforms(fs["model"])
    ctx.shadowmaps()
    sms = [ctx.shadow(i) for i in range(len(fs["lights"]))]
    ctx.warpmap(fs["vox"], _warp_params())
    cap = {"occ": ctx.warp_occupancy(), "wm": ctx.warp_texture(), "shadow": sms}
    assert fs["vox"].warp_texture and fs["inj"].warp_texture and fs["trace"].warp_texture
    ctx.voxelize(fs["vox"])
    cap["albedo"], cap["normal"] = ctx.albedo(), ctx.normal()
    ctx.normalize_inject(fs["inj"])
    cap["rad0"], cap["litmask"] = ctx.radiance(0), ctx.litmask()
    ctx.mip(fs["mip"])
    cap["levels"] = [ctx.radiance(l) for l in range(g.levels)]
    cap["faces"] = [None] + [[ctx.radiance(l, f) for f in range(6)] for l in range(1, g.levels)]
    ctx.gbuffer(fs["cam"])
    cam = fs["cam"]
    cap["gbuf"] = ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(cam.height, cam.width, 16)
    ctx.trace(fs["trace"])
    cap["indirect"] = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
    cap["specular"] = ctx.download(abi.VCT_BUF_SPECULAR, 0, np.float32).reshape(-1, 4)
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    yield app, fs, g, cap, sc
    app.close()


def test_scene5_shadow_maps_bit_exact(reduced):
    app, fs, g, cap, sc = reduced
    for i, L in enumerate(fs["lights"]):
        ref = ob.shadowmap(sc, L)
        assert np.array_equal(cap["shadow"][i].view(np.uint32), ref.view(np.uint32)), f"light {i}"
    assert sum((s < 1).mean() > 0.01 for s in cap["shadow"]) >= 4  # most maps see geometry


def test_scene5_warp_texture_bit_exact(reduced):
    app, fs, g, cap, sc = reduced
    occ = ob.warp_occupancy(sc, fs["vox"])
    assert np.array_equal(cap["occ"], occ)
    wm = ob.warpmap(occ, ob.warp_params(high=4.0))
    assert np.array_equal(cap["wm"].view(np.uint32), wm.view(np.uint32))
    assert not np.array_equal(wm, ob.warpmap(occ))  # the high resolution 4.0 shapes the map


def test_scene5_warped_voxelize_inject_and_fp8_pyramid_bit_exact(reduced):
    app, fs, g, cap, sc = reduced
    wm = cap["wm"]
    alb, nrm, _ = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim, warpmap=wm)
    assert np.array_equal(cap["albedo"], alb), f"{(cap['albedo'] != alb).sum()} voxels differ"
    assert np.array_equal(cap["normal"], nrm)
    _, rad0, lm, st = ob.normalize_inject(alb, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), fs["lights"],
                                          cap["shadow"], warpmap=wm)
    assert np.array_equal(cap["litmask"], lm)
    assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))
    assert st.lit_voxels > 0
    faces = ob.mip_aniso(cap["rad0"], g.levels)
    for l in range(1, g.levels):
        for f in range(6):
            r = faces[l][f] if l < 3 else ob.fp8_quantize(faces[l][f])
            assert np.array_equal(cap["faces"][l][f].view(np.uint32), r.view(np.uint32)), (l, f)


def test_scene5_warped_trace_within_tolerance(reduced):
    app, fs, g, cap, sc = reduced
    gref = ob.gbuffer(sc, fs["cam"])
    assert np.array_equal(cap["gbuf"].view(np.uint32), gref.view(np.uint32))
    tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                         app.mesh()["rough"] is not None, 1, warpmap=np.ascontiguousarray(cap["wm"]))
    ind, spec = ob.trace(cap["gbuf"], cap["levels"], tp, aniso_faces=cap["faces"])
    ei, es = rel_err(cap["indirect"][:, :3], ind[:, :3]), rel_err(cap["specular"], spec)
    print(f"scene 5 @64^3: warped indirect {ei:.3e}, specular {es:.3e} vs the oracle", flush=True)
    assert ei < 1e-3 and es < 1e-3
    assert rel_err(cap["indirect"][:, 3], ind[:, 3]) < 1e-3
    assert ind[:, :3].sum() > 0 and spec[:, :3].sum() > 0


@pytest.fixture(scope="module")
def full():
    app = make_app(5)
    yield app
    app.close()


def _grab(app):
    c = app.ctx
    return dict(alb=c.albedo(), rad0=c.radiance(0), l4=c.radiance(4, 3), wm=c.warp_texture(),
                occ=c.warp_occupancy(), ind=c.download(abi.VCT_BUF_INDIRECT, 0, np.float32),
                spec=c.download(abi.VCT_BUF_SPECULAR, 0, np.float32), stats=c.stats())


def test_scene5_full_size_frame_is_deterministic(full):
    print("scene 5 full size: frames 3, 4, 3", flush=True)
    full.render(3)
    a = _grab(full)
    full.render(4)
    full.render(3)
    b = _grab(full)
    for k in ("alb", "rad0", "l4", "wm", "occ", "ind", "spec"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    s = a["stats"]
    assert s.unique_voxels == int(((a["alb"] >> 24) > 0).sum())
    assert s.total_voxel_fragments >= s.unique_voxels > 100_000
    assert 0 < s.lit_voxels <= s.unique_voxels
    ind = a["ind"].reshape(-1, 4)
    assert np.isfinite(ind).all() and ind[:, :3].sum() > 0
    assert ind.shape[0] == 3840 * 2160


def test_scene5_full_size_warp_texture_follows_the_rule(full):
    full.render(6)
    occ, wm = full.ctx.warp_occupancy(), full.ctx.warp_texture()
    n = int((occ > 0).sum())
    assert 0 < n < 32 ** 3
    ref = ob.warpmap(occ, ob.warp_params(high=4.0))
    assert np.array_equal(wm.view(np.uint32), ref.view(np.uint32))
    assert np.all(np.diff(wm[..., 0], axis=2) > 0) and np.all(np.diff(wm[..., 1], axis=1) > 0)
    assert np.all(np.diff(wm[..., 2], axis=0) > 0)
    assert wm[..., :3].min() > 0 and wm[..., :3].max() < 1

"""GPU parity of the warp texture path (warpTexture; src/Application.cpp:235-575,
generateWarpmap.frag, common.glsl:48-50, voxelize.frag:103-105,
phong.frag:150-152): the 32^3 occupancy grid and warp texture bit-exact, the
warped voxelization and injection bit-exact, the warped trace within the
1e-3 mean relative error of the indirect image, on the small scene and on the
dynamic scene at reduced size."""
import numpy as np
import pytest

from gpu_util import make_app, rel_err
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(1, {}), (2, dict(dim=128, levels=7, width=480, height=270,
                                                          shadow_size=256))], ids=["scene1", "scene2_128"])
def warped(request):
    scene, kw = request.param
    app = make_app(scene, **kw)
    ctx = app.ctx
    fs = app.frame_state(2)
    g = ctx.grid()
    ctx.set_transforms(fs["model"])
    sms = []
    for i, L in enumerate(fs["lights"]):
        if L.shadow_caster:
            ctx.shadowmap(i)
            sms.append(ctx.shadow(i))
        else:
            sms.append(None)
    print(f"scene {scene}: warpmap", flush=True)
    ctx.warpmap(fs["vox"])
    cap = {"occ": ctx.warp_occupancy(), "wm": ctx.warp_texture(), "shadow": sms}
    vox = abi.VoxelizeParams.from_buffer_copy(fs["vox"])
    vox.warp_texture = 1
    ctx.voxelize(vox)
    cap["albedo"], cap["normal"] = ctx.albedo(), ctx.normal()
    inj = abi.InjectParams.from_buffer_copy(fs["inj"])
    inj.warp_texture = 1
    ctx.normalize_inject(inj)
    cap["rad0"], cap["litmask"] = ctx.radiance(0), ctx.litmask()
    ctx.mip(fs["mip"])
    cap["levels"] = [ctx.radiance(l) for l in range(g.levels)]
    if g.anisotropic:
        cap["faces"] = [None] + [[ctx.radiance(l, f) for f in range(6)] for l in range(1, g.levels)]
    ctx.gbuffer(fs["cam"])
    cam = fs["cam"]
    cap["gbuf"] = ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(cam.height, cam.width, 16)
    tr = abi.TraceParams.from_buffer_copy(fs["trace"])
    tr.warp_texture = 1
    ctx.trace(tr)
    cap["indirect"] = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
    if tr.enable_reflections:
        cap["specular"] = ctx.download(abi.VCT_BUF_SPECULAR, 0, np.float32).reshape(-1, 4)
    yield scene, app, fs, g, cap, tr
    app.close()


def test_occupancy_and_warp_texture_bit_exact(warped):
    scene, app, fs, g, cap, _ = warped
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    occ = ob.warp_occupancy(sc, fs["vox"])
    assert np.array_equal(cap["occ"], occ), f"{(cap['occ'] != occ).sum()} occupancy cells differ"
    wm = ob.warpmap(occ)
    assert np.array_equal(cap["wm"].view(np.uint32), wm.view(np.uint32))
    n = int((occ > 0).sum())
    print(f"scene {scene}: {n} of 32768 warp cells occupied", flush=True)
    assert 0 < n < 32 ** 3


def test_warped_voxelize_and_inject_bit_exact(warped):
    scene, app, fs, g, cap, _ = warped
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    wm = ob.warpmap(ob.warp_occupancy(sc, fs["vox"]))
    alb, nrm, _ = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim, warpmap=wm)
    assert np.array_equal(cap["albedo"], alb), f"{(cap['albedo'] != alb).sum()} voxels differ"
    assert np.array_equal(cap["normal"], nrm)
    plain, _, _ = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim)
    assert not np.array_equal(plain, alb)  # the warp moved voxels
    _, rad0, lm, _ = ob.normalize_inject(alb, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), fs["lights"],
                                         cap["shadow"], warpmap=wm)
    assert np.array_equal(cap["litmask"], lm)
    assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))


def test_warped_trace_within_tolerance(warped):
    scene, app, fs, g, cap, tr = warped
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    wm = np.ascontiguousarray(ob.warpmap(ob.warp_occupancy(sc, fs["vox"])))
    tp = ob.trace_params(tr, g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                         app.mesh()["rough"] is not None, 1 if g.anisotropic else 0, warpmap=wm)
    ind, spec = ob.trace(cap["gbuf"], cap["levels"], tp, aniso_faces=cap.get("faces"))
    e = rel_err(cap["indirect"][:, :3], ind[:, :3])
    print(f"scene {scene}: warped indirect mean relative error vs the FP32 oracle {e:.3e}", flush=True)
    assert e < 1e-3
    assert rel_err(cap["indirect"][:, 3], ind[:, 3]) < 1e-3
    if "specular" in cap:
        assert rel_err(cap["specular"], spec) < 1e-3
    assert ind[:, :3].sum() > 0


def test_application_frame_with_warp_texture():
    app = make_app(1, warp_texture=1)
    try:
        app.render(0)
        pt = app.ctx.pass_times()
        assert pt.warpmap_ms > 0
        col = app.ctx.color()
        assert np.isfinite(col).all() and col[:, :3].sum() > 0
    finally:
        app.close()

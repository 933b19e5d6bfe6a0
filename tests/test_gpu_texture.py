"""GPU parity of the diffuse maps (vct_set_textures, texture(diffuseMap, uv) in
voxelize.frag:197, phong.frag:449 and testTesselation.tese:122) against the
oracle on the same textured scene: the voxelizer's albedo grid (lit and
unlit), the G-buffer (its albedo the sampled map, as RGBA8) and the
tessellation voxelizer bit-exact.  The maps cover the sampler's cases: a full
mip chain, a non-square complete chain, an incomplete chain and an index
outside the table (both (0, 0, 0, 1)), and untextured triangles; then a map
decoded from a DDS file and an OBJ whose material names it."""
import numpy as np
import pytest

import vct_amd
from dds_fixtures import random_dds
from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def rand_texture(rng, W, H, levels):
    n = sum(max(1, W >> l) * max(1, H >> l) for l in range(levels))
    return (W, H, levels, rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32))


def textured(mesh, rng, textures):
    p = mesh["pos"]
    uv = np.stack([0.7 * p[..., 0] + 0.2 * p[..., 1], 0.6 * p[..., 2] - 0.3 * p[..., 1] + 0.1], -1)
    m = dict(mesh)
    m.update(uv=uv.astype(np.float32), textures=textures,
             texid=rng.choice([-1, 0, 0, 0, 1, 2, 7], size=len(mesh["albedo"])).astype(np.int32))
    return m


def upload(ctx, m):
    ctx.set_scene(m["pos"], m["nrm"], m["tan"], m["albedo"], m["obj"], m["nobj"], m["rough"], m["vtan"],
                  m["uv"], m["texid"])
    ctx.set_textures(m["textures"])


def check_passes(app, m, frame, label):
    ctx = app.ctx
    fs = app.frame_state(frame)
    ctx.set_transforms(fs["model"])
    sc = ob.SceneArrays(m, fs["model"])
    g = ctx.grid()
    sms = []
    for i, L in enumerate(fs["lights"]):
        if L.shadow_caster:
            ctx.shadowmap(i)
            sms.append(ctx.shadow(i))
        else:
            sms.append(None)
    vox = fs["vox"]
    vox.atomic_max = abi.VCT_ACCUM_MAX
    for lighting in (0, 1):
        print(f"textures ({label}): voxelize, lighting {lighting}", flush=True)
        vox.lighting = lighting
        ctx.voxelize(vox)
        alb = ctx.albedo()
        ralb, rnrm, st = ob.voxelize(sc, fs["lights"], sms, vox, g.dim)
        assert (ralb != 0).sum() > 100
        assert np.array_equal(alb, ralb), f"{(alb != ralb).sum()} albedo voxels differ"
        assert ctx.stats().total_voxel_fragments == st.total_voxel_fragments
    print(f"textures ({label}): G-buffer", flush=True)
    cam = fs["cam"]
    ctx.gbuffer(cam)
    got = ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(cam.height, cam.width, 16)
    ref = ob.gbuffer(sc, cam)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"{(got.view(np.uint32) != ref.view(np.uint32)).any(-1).sum()} G-buffer texels differ"
    tri = got[..., 15].view(np.uint32)[got[..., 11].view(np.uint32) == 1]
    shown = m["texid"][tri]
    assert (shown >= 0).mean() > 0.2  # textured surfaces are in view
    print(f"textures ({label}): tessellation voxelizer", flush=True)
    ctx.voxelize_tessellated(vox)
    ralb, rnrm = ob.voxelize_tess(sc, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), accum=1)
    assert np.array_equal(ctx.albedo(), ralb)
    return alb


@pytest.mark.parametrize("scene,kw", [(1, {}), (2, dict(dim=128, levels=7, width=256, height=144, shadow_size=256))])
def test_textured_scene_matches_oracle(scene, kw):
    app = make_app(scene, **kw)
    try:
        rng = np.random.default_rng(10 + scene)
        textures = [rand_texture(rng, 64, 64, 7), rand_texture(rng, 32, 8, 6), rand_texture(rng, 16, 16, 3)]
        m = textured(app.mesh(), rng, textures)
        upload(app.ctx, m)
        alb = check_passes(app, m, 2, f"scene {scene}")
        # the maps change the grid: the same scene without them differs
        plain = dict(m, texid=np.full_like(m["texid"], -1))
        upload(app.ctx, plain)
        fs = app.frame_state(2)
        app.ctx.set_transforms(fs["model"])
        for i, L in enumerate(fs["lights"]):
            if L.shadow_caster:
                app.ctx.shadowmap(i)
        app.ctx.voxelize(fs["vox"])
        assert (app.ctx.albedo() != alb).sum() > 100
    finally:
        app.close()


def test_dds_map_matches_oracle(tmp_path):
    rng = np.random.default_rng(3)
    p = tmp_path / "r.dds"
    p.write_bytes(random_dds(rng, 32, 32, b"DXT5"))
    tex = vct_amd.load_dds(p)
    assert tex[2] == 6
    app = make_app(1)
    try:
        m = textured(app.mesh(), rng, [tex])
        m["texid"] = np.where(m["texid"] >= 0, 0, -1).astype(np.int32)
        upload(app.ctx, m)
        check_passes(app, m, 0, "DDS")
    finally:
        app.close()


def test_obj_with_a_dds_map_renders_it(tmp_path):
    from obj_fixtures import uv_sphere_obj
    rng = np.random.default_rng(4)
    (tmp_path / "r.dds").write_bytes(random_dds(rng, 64, 64, b"DXT1"))
    (tmp_path / "s.mtl").write_text("newmtl m\nKd 1 1 1\nmap_Kd r.dds\n")
    (tmp_path / "s.obj").write_text("mtllib s.mtl\nusemtl m\n" + uv_sphere_obj(16, 32))
    app = vct_amd.App(obj_path=tmp_path / "s.obj", device=0, width=256, height=192)
    try:
        print("textures: OBJ frame", flush=True)
        app.render(0)
        loaded = vct_amd.load_obj(tmp_path / "s.obj")
        assert (loaded["texid"] == 0).all() and len(loaded["textures"]) == 1
        m = app.mesh()
        m.update(uv=loaded["uv"], texid=loaded["texid"], textures=loaded["textures"])
        fs = app.frame_state(0)
        cam = fs["cam"]
        app.ctx.gbuffer(cam)
        got = app.ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(cam.height, cam.width, 16)
        ref = ob.gbuffer(ob.SceneArrays(m, fs["model"]), cam)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        cov = got[..., 11].view(np.uint32) == 1
        albs = np.unique(got[..., 3].view(np.uint32)[cov])
        assert cov.mean() > 0.1 and albs.size > 50  # the map, not one Kd colour
    finally:
        app.close()

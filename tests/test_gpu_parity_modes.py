"""GPU parity beyond the small scene: the dynamic scene at reduced resolution
(5 shadow-casting lights including perspective point-light maps, 6-direction
mips, the specular cone, the composed image with point lights and reflections),
the FP32 and FP8 storage formats on isotropic and 6-direction pyramids, and
fill-holes."""
import numpy as np
import pytest

from gpu_util import (assert_compose_close, compose_reference, compose_variant, make_app, rel_err,
                      run_frame_capture)
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene2_small():
    app = make_app(2, dim=128, levels=7, width=480, height=270, shadow_size=256)
    cap = run_frame_capture(app, 5)
    yield app, cap
    app.close()


def test_scene2_shadow_maps_bit_exact(scene2_small):
    app, cap = scene2_small
    fs = cap["fs"]
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    for i, L in enumerate(fs["lights"]):
        ref = ob.shadowmap(sc, L)
        assert np.array_equal(cap["shadow"][i].view(np.uint32), ref.view(np.uint32)), f"light {i}"
        assert (ref < 1).mean() > 0.01  # the map sees geometry


def test_scene2_voxelize_and_inject_bit_exact(scene2_small):
    app, cap = scene2_small
    fs, g = cap["fs"], cap["grid"]
    sc = ob.SceneArrays(app.mesh(), fs["model"])
    alb, nrm, st = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim)
    assert np.array_equal(cap["albedo_vox"], alb), f"{(cap['albedo_vox'] != alb).sum()} voxels differ"
    assert np.array_equal(cap["normal_vox"], nrm)
    assert cap["stats_vox"].total_voxel_fragments == st.total_voxel_fragments
    a2, rad0, lm, st2 = ob.normalize_inject(alb, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                                            fs["lights"], cap["shadow"])
    assert np.array_equal(cap["litmask"], lm)
    assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))
    assert cap["stats"].lit_voxels == st2.lit_voxels > 0


def test_scene2_aniso_mips_bit_exact(scene2_small):
    app, cap = scene2_small
    g = cap["grid"]
    faces = ob.mip_aniso(cap["rad0"], g.levels)
    for l in range(1, g.levels):
        for f in range(6):
            got = cap["faces"][l][f]
            assert np.array_equal(got.view(np.uint32), faces[l][f].view(np.uint32)), (l, f)


def test_scene2_albedo_chain_bit_exact(scene2_small):
    """The voxelColor chain the mipmap pass filters beside the radiance
    (Application.cpp:904-917), from the normalized albedo grid."""
    app, cap = scene2_small
    g = cap["grid"]
    ref = ob.mip_albedo(cap["albedo_norm"], g.levels)
    for l in range(1, g.levels):
        got = cap["albedo_levels"][l]
        assert np.array_equal(got, ref[l]), (l, int((got != ref[l]).sum()))
    assert (ref[1] >> 24).any()


def test_scene2_gbuffer_and_trace(scene2_small):
    app, cap = scene2_small
    fs, g = cap["fs"], cap["grid"]
    mesh = app.mesh()
    sc = ob.SceneArrays(mesh, fs["model"])
    gref = ob.gbuffer(sc, fs["cam"])
    assert np.array_equal(cap["gbuf"].view(np.uint32), gref.view(np.uint32))
    levels = cap["levels"]
    tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                         mesh["rough"] is not None, 1)
    ind, spec = ob.trace(cap["gbuf"], levels, tp, aniso_faces=cap["faces"])
    ei, es = rel_err(cap["indirect"][:, :3], ind[:, :3]), rel_err(cap["specular"], spec)
    print(f"scene 2 mean relative error vs the FP32 oracle: indirect {ei:.3e}, specular {es:.3e}")
    assert ei < 1e-3
    assert rel_err(cap["indirect"][:, 3], ind[:, 3]) < 1e-3
    assert es < 1e-3
    assert ind[:, :3].sum() > 0 and spec[:, :3].sum() > 0


@pytest.mark.parametrize("variant", [dict(), dict(cook_torrance=0), dict(enable_reflections=0)],
                         ids=["app", "blinn_phong", "no_reflections"])
def test_scene2_compose(scene2_small, variant):
    app, cap = scene2_small
    p = compose_variant(cap["compose_params"], **variant)
    assert cap["compose_params"].enable_reflections == 1
    if variant:
        app.ctx.compose(p)
        got = app.ctx.color()
    else:
        got = cap["color"]
    ref = compose_reference(app, cap, p)
    assert_compose_close(got, ref, f"scene 2 compose {variant}")
    assert ref[:, :3].sum() > 0


@pytest.mark.parametrize("storage", [1, 2])
def test_storage_formats(storage):
    app = make_app(1, storage=storage)
    cap = run_frame_capture(app, 0)
    g, fs = cap["grid"], cap["fs"]
    ref = ob.mip(cap["rad0"], g.levels, fp32=(storage == 1))
    for l in range(1, g.levels):
        r = ref[l]
        if storage == 2 and l >= 3:
            r = ob.fp8_quantize(r)  # FP8 E4M3 with a per-4^3-brick fp32 scale
        got = cap["levels"][l]
        assert np.array_equal(got.view(np.uint32), r.view(np.uint32)), (storage, l, np.abs(got - r).max())
    mesh = app.mesh()
    tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                         mesh["rough"] is not None)
    ind, _ = ob.trace(cap["gbuf"], cap["levels"], tp)
    assert rel_err(cap["indirect"][:, :3], ind[:, :3]) < 1e-3
    app.close()


@pytest.mark.parametrize("storage", [1, 2])
def test_scene2_storage_formats_aniso(storage):
    """FP32 and FP8 (+ per-(brick, face) scale) storage of the 6-direction
    pyramid: level 0, every face of every level bit for bit, and the traced
    indirect and specular images against the oracle on the decoded faces."""
    app = make_app(2, dim=128, levels=7, width=480, height=270, shadow_size=256, storage=storage)
    try:
        print(f"scene 2 @128^3, anisotropic, storage {storage}", flush=True)
        cap = run_frame_capture(app, 5)
        fs, g = cap["fs"], cap["grid"]
        assert g.anisotropic
        fp32 = storage == 1
        _, rad0, _, _ = ob.normalize_inject(cap["albedo_vox"], g.dim, tuple(g.center), tuple(g.vmin),
                                            tuple(g.vmax), fs["lights"], cap["shadow"], fp32=fp32)
        assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))
        faces = ob.mip_aniso(cap["rad0"], g.levels, fp32=fp32)
        for l in range(1, g.levels):
            for f in range(6):
                r = faces[l][f]
                if storage == 2 and l >= 3:
                    r = ob.fp8_quantize(r)  # E4M3 with an fp32 scale per 4^3 brick of this face
                got = cap["faces"][l][f]
                assert np.array_equal(got.view(np.uint32), r.view(np.uint32)), (storage, l, f,
                                                                                np.abs(got - r).max())
        mesh = app.mesh()
        tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                             mesh["rough"] is not None, 1)
        ind, spec = ob.trace(cap["gbuf"], cap["levels"], tp, aniso_faces=cap["faces"])
        ei, es = rel_err(cap["indirect"][:, :3], ind[:, :3]), rel_err(cap["specular"], spec)
        print(f"storage {storage}: indirect {ei:.3e}, specular {es:.3e} vs the oracle", flush=True)
        assert ei < 1e-3 and es < 1e-3
        assert rel_err(cap["indirect"][:, 3], ind[:, 3]) < 1e-3
        assert ind[:, :3].sum() > 0 and spec[:, :3].sum() > 0
    finally:
        app.close()


def test_fill_holes_bit_exact():
    app = make_app(1, fill_holes=1)
    cap = run_frame_capture(app, 0, fill_holes=True)
    ref = ob.fill_holes(cap["rad0"])
    assert np.array_equal(cap["rad0_filled"].view(np.uint32), ref.view(np.uint32))
    assert (ref[..., 3] > 0).sum() > (cap["rad0"][..., 3] > 0).sum()
    lv = ob.mip(ref, cap["grid"].levels)
    for l in range(1, cap["grid"].levels):
        assert np.array_equal(cap["levels"][l].view(np.uint32), lv[l].view(np.uint32))
    app.close()

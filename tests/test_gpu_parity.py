"""GPU parity: each pass of the HIP path against the oracle on the same inputs.

Bars (north-star): voxel occupancy, the packed RGBA8 albedo/normal
accumulators and the statistics bit-exact; shadow depth, radiance level 0 and
the fp16 mip chain bit-exact (identical operation order; the required
tolerance would be 1e-3 max-abs); the indirect-lighting image within 1e-3 mean
relative error of the FP32 oracle; the composed image within 1e-5 relative per
channel (same operation order, powf/sqrtf of two math libraries).
"""
import numpy as np
import pytest

from gpu_util import (assert_compose_close, compose_reference, compose_variant, make_app, rel_err,
                      run_frame_capture)
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene1():
    app = make_app(1)
    cap = run_frame_capture(app, 0, fill_holes=False)
    yield app, cap
    app.close()


def _oracle_scene(app, fs):
    return ob.SceneArrays(app.mesh(), fs["model"])


def test_shadowmap_bit_exact(scene1):
    app, cap = scene1
    fs = cap["fs"]
    sc = _oracle_scene(app, fs)
    ref = ob.shadowmap(sc, fs["lights"][0])
    got = cap["shadow"][0]
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"{(got != ref).sum()} shadow texels differ"


def test_voxelize_bit_exact(scene1):
    app, cap = scene1
    fs, g = cap["fs"], cap["grid"]
    sc = _oracle_scene(app, fs)
    alb, nrm, st = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim)
    occ_g = (cap["albedo_vox"] >> 24) > 0
    occ_r = (alb >> 24) > 0
    assert np.array_equal(occ_g, occ_r), f"occupancy differs in {(occ_g != occ_r).sum()} voxels"
    assert np.array_equal(cap["albedo_vox"], alb), f"albedo differs in {(cap['albedo_vox'] != alb).sum()} voxels"
    assert np.array_equal(cap["normal_vox"], nrm), f"normal differs in {(cap['normal_vox'] != nrm).sum()} voxels"
    assert cap["stats_vox"].total_voxel_fragments == st.total_voxel_fragments


def test_normalize_inject_bit_exact(scene1):
    app, cap = scene1
    fs, g = cap["fs"], cap["grid"]
    alb, rad0, lm, st = ob.normalize_inject(cap["albedo_vox"], g.dim, tuple(g.center), tuple(g.vmin),
                                           tuple(g.vmax), fs["lights"], cap["shadow"])
    assert np.array_equal(cap["litmask"], lm)
    assert np.array_equal(cap["albedo_norm"], alb)
    assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))
    s = cap["stats"]
    assert (s.unique_voxels, s.max_fragments_per_voxel, s.lit_voxels) == \
        (st.unique_voxels, st.max_fragments_per_voxel, st.lit_voxels)
    # uniqueVoxels is the popcount of the occupancy
    assert s.unique_voxels == int(((cap["albedo_vox"] >> 24) > 0).sum())


def test_mip_chain_bit_exact(scene1):
    app, cap = scene1
    g = cap["grid"]
    ref = ob.mip(cap["rad0"], g.levels)
    for l in range(1, g.levels):
        got = cap["levels"][l]
        assert np.array_equal(got.view(np.uint32), ref[l].view(np.uint32)), \
            f"level {l}: max abs {np.abs(got - ref[l]).max()}"


def test_albedo_chain_bit_exact(scene1):
    app, cap = scene1
    g = cap["grid"]
    ref = ob.mip_albedo(cap["albedo_norm"], g.levels)
    for l in range(1, g.levels):
        assert np.array_equal(cap["albedo_levels"][l], ref[l]), f"albedo level {l}"


def test_gbuffer_matches_oracle(scene1):
    app, cap = scene1
    fs = cap["fs"]
    sc = _oracle_scene(app, fs)
    ref = ob.gbuffer(sc, fs["cam"])
    got = cap["gbuf"]
    tri_g = got[..., 15].view(np.uint32)
    tri_r = ref[..., 15].view(np.uint32)
    cov_g = got[..., 11].view(np.uint32)
    assert (cov_g == ref[..., 11].view(np.uint32)).all()
    assert np.array_equal(tri_g, tri_r), f"visible triangle differs at {(tri_g != tri_r).sum()} pixels"
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
        f"attributes max abs diff {np.abs(got - ref).max()}"


def test_trace_within_tolerance(scene1):
    app, cap = scene1
    fs, g = cap["fs"], cap["grid"]
    mesh = app.mesh()
    tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                         mesh["rough"] is not None)
    ref, _ = ob.trace(cap["gbuf"], cap["levels"], tp)
    got = cap["indirect"]
    e = rel_err(got[:, :3], ref[:, :3])
    print(f"indirect image mean relative error vs the FP32 oracle: {e:.3e}")
    assert e < 1e-3, f"indirect image mean relative error {e}"
    assert rel_err(got[:, 3], ref[:, 3]) < 1e-3
    assert ref[:, :3].sum() > 0  # the scene is lit: the comparison is not vacuous


@pytest.mark.parametrize("variant", [dict(), dict(cook_torrance=0), dict(postprocess=0),
                                     dict(enable_indirect=0), dict(enable_diffuse=0, enable_specular=0)],
                         ids=["app", "blinn_phong", "no_post", "no_indirect", "indirect_only"])
def test_compose_matches_oracle(scene1, variant):
    app, cap = scene1
    p = compose_variant(cap["compose_params"], **variant)
    if variant:
        app.ctx.compose(p)
        got = app.ctx.color()
    else:
        got = cap["color"]
    ref = compose_reference(app, cap, p)
    assert_compose_close(got, ref, f"compose {variant}")
    cov = cap["gbuf"][..., 11].view(np.uint32).reshape(-1) != 0
    assert ref[cov, :3].sum() > 0 and np.all(ref[~cov] == 0)


def test_normalize_inject_alone_bit_exact():
    """A one-level grid (no mip chain at all): the injection alone matches the
    oracle, statistics included."""
    app = make_app(1, levels=1)
    try:
        ctx = app.ctx
        g = ctx.grid()
        assert g.levels == 1
        fs = app.frame_state(0)
        ctx.set_transforms(fs["model"])
        sms = []
        for i, L in enumerate(fs["lights"]):
            if L.shadow_caster:
                ctx.shadowmap(i)
                sms.append(ctx.shadow(i))
            else:
                sms.append(None)
        ctx.voxelize(fs["vox"])
        alb_vox = ctx.albedo()
        ctx.normalize_inject(fs["inj"])
        alb, rad0, lm, st = ob.normalize_inject(alb_vox, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                                               fs["lights"], sms)
        assert np.array_equal(ctx.litmask(), lm)
        assert np.array_equal(ctx.albedo(), alb)
        assert np.array_equal(ctx.radiance(0).view(np.uint32), rad0.view(np.uint32))
        s = ctx.stats()
        assert (s.unique_voxels, s.max_fragments_per_voxel, s.lit_voxels) == \
            (st.unique_voxels, st.max_fragments_per_voxel, st.lit_voxels)
        assert st.lit_voxels > 0
    finally:
        app.close()

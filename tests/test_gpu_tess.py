"""GPU parity of the tessellation voxelizer (vct_voxelize_tessellated;
simpleTesselated.vert + testTesselation.tesc/.tese, Application.cpp:585-654)
against the oracle's or_voxelize_tess on the same scene: bit-exact albedo and
normal grids with atomicMax, and for the RGBA8 running average (whose rgb
depends on the order of the points) the same occupied voxels and the same
point count in every voxel's alpha byte.  Then a whole frame through the app
with the flag set."""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene,kw", [(1, {}), (2, dict(dim=128, levels=7, width=256, height=144, shadow_size=256))])
def test_tessellated_voxels_match_oracle(scene, kw):
    app = make_app(scene, **kw)
    try:
        ctx = app.ctx
        fs = app.frame_state(3)
        g = ctx.grid()
        ctx.set_transforms(fs["model"])
        vox = fs["vox"]
        sc = ob.SceneArrays(app.mesh(), fs["model"])
        print(f"tessellation voxelizer: scene {scene}, {g.dim}^3, atomicMax", flush=True)
        vox.atomic_max = abi.VCT_ACCUM_MAX
        ctx.voxelize_tessellated(vox)
        alb, nrm = ctx.albedo(), ctx.normal()
        ralb, rnrm = ob.voxelize_tess(sc, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), accum=1)
        assert (ralb != 0).sum() > 100
        assert np.array_equal(alb, ralb), f"{(alb != ralb).sum()} albedo voxels differ"
        assert np.array_equal(nrm, rnrm), f"{(nrm != rnrm).sum()} normal voxels differ"
        assert ctx.stats().total_voxel_fragments == 0
        print(f"tessellation voxelizer: scene {scene}, running average", flush=True)
        vox.atomic_max = abi.VCT_ACCUM_AVG
        ctx.voxelize_tessellated(vox)
        alb2 = ctx.albedo()
        ralb2, _ = ob.voxelize_tess(sc, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), accum=0)
        assert np.array_equal(alb2 != 0, ralb2 != 0)
        # the alpha byte counts the voxel's points (mod 256) whatever their order
        assert np.array_equal(alb2 >> 24, ralb2 >> 24)
    finally:
        app.close()


def test_app_frame_with_the_tessellation_voxelizer():
    app = make_app(1, tessellation=1)
    try:
        print("tessellation voxelizer: app frame", flush=True)
        app.render(0)
        ind = app.ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
        assert np.isfinite(ind).all() and ind[:, :3].sum() > 0
        assert app.ctx.stats().unique_voxels > 0
    finally:
        app.close()

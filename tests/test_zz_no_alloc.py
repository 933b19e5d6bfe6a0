"""vct_normalize_inject_noalloc against vct_normalize_inject: the same radiance
level 0 (raw stored bytes), albedo opacity rewrite and statistics, through the
C ABI, on two calls in a row -- a 128^3 grid, then a smaller 64^3 one (the
second call reuses the static scratch the first one filled) -- with and
without the temporal filter, and the size limit."""
import numpy as np
import pytest

from gpu_util import make_app
from vct_amd import VCTError
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def _state(ctx):
    raw = ctx.download(abi.VCT_BUF_RADIANCE_RAW, 0, np.uint8)
    st = ctx.stats()
    return raw, ctx.albedo(), (st.unique_voxels, st.max_fragments_per_voxel, st.lit_voxels)


def _run(app, inj, alb_vox, prev0, noalloc):
    """Level 0 := prev0 (the previous frame for the temporal filter, garbage
    otherwise, so a call that wrote nothing would show), albedo := the
    voxelized grid, then one injection."""
    ctx = app.ctx
    ctx.upload(abi.VCT_BUF_ALBEDO, 0, alb_vox)
    ctx.upload(abi.VCT_BUF_RADIANCE, 0, prev0)
    if noalloc:
        ctx.normalize_inject_noalloc(inj)
    else:
        ctx.normalize_inject(inj)
    return _state(ctx)


def _compare(scene, dim, temporal):
    app = make_app(scene, dim=dim, levels=7)
    try:
        ctx = app.ctx
        fs = app.frame_state(0)
        inj = abi.InjectParams.from_buffer_copy(fs["inj"])
        inj.temporal_filter, inj.temporal_decay = (1 if temporal else 0), 0.8
        print(f"no-alloc inject: scene {scene}, {dim}^3, temporal {temporal}", flush=True)
        ctx.set_transforms(fs["model"])
        ctx.shadowmaps()
        ctx.voxelize(fs["vox"])
        alb_vox = ctx.albedo()
        rng = np.random.default_rng(dim)
        prev0 = (rng.integers(0, 64, size=(dim, dim, dim, 4)) / 64.0).astype(np.float32)
        # the no-alloc call first, on a context whose statistics are still zero
        got = _run(app, inj, alb_vox, prev0, noalloc=True)
        ref = _run(app, inj, alb_vox, prev0, noalloc=False)
        assert ref[2][0] > 0 and ref[2][2] > 0  # occupied and lit voxels
        assert np.array_equal(got[0], ref[0]), "radiance level 0 bytes"
        assert np.array_equal(got[1], ref[1]), "albedo opacity rewrite"
        assert got[2] == ref[2], (got[2], ref[2])
    finally:
        app.close()


@pytest.mark.parametrize("temporal", [False, True])
def test_noalloc_matches_plain_then_smaller_grid(temporal):
    _compare(2, 128, temporal)
    _compare(1, 64, temporal)


def test_noalloc_rejects_grids_beyond_the_scratch():
    import vct_amd
    ctx = vct_amd.Context(dim=512, levels=2, device=0)
    try:
        with pytest.raises(VCTError) as e:
            ctx.normalize_inject_noalloc(abi.InjectParams())
        assert "256" in str(e.value)
    finally:
        ctx.close()

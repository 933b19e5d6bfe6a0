"""The voxelizer's other accumulation modes against the oracle.

  VCT_ACCUM_AVG  imageAtomicRGBA8Avg, the reference's RGBA8 path with
                 voxelizeAtomicMax off (voxelize.frag:111-139): the alpha byte
                 counts a voxel's fragments (mod 256) whatever their order, so
                 it is compared bit for bit; the running-average colour depends
                 on the order in which fragments land, so only voxels with one
                 fragment, whose average is that fragment, are compared bit for
                 bit (the oracle's own average is one order of many).
  VCT_ACCUM_SUM  fixed-point sums and normalizeVoxels.comp: exact integers, so
                 the normalized albedo, normal, radiance level 0 and statistics
                 are compared bit for bit."""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu

CASES = [(1, {}), (2, dict(dim=128, levels=7, width=64, height=64, shadow_size=256))]


def _setup(scene, kw):
    app = make_app(scene, **kw)
    ctx = app.ctx
    fs = app.frame_state(1)
    ctx.set_transforms(fs["model"])
    sms = []
    for i, L in enumerate(fs["lights"]):
        if L.shadow_caster:
            ctx.shadowmap(i)
            sms.append(ctx.shadow(i))
        else:
            sms.append(None)
    return app, ctx, fs, sms


def _bytes(a):
    return np.stack([(a >> (8 * k)) & 255 for k in range(4)], -1).astype(np.int64)


@pytest.mark.parametrize("scene,kw", CASES, ids=["scene1", "scene2_128"])
def test_running_average_mode(scene, kw):
    app, ctx, fs, sms = _setup(scene, kw)
    try:
        g = ctx.grid()
        print(f"scene {scene}: running-average voxelization", flush=True)
        vox = abi.VoxelizeParams.from_buffer_copy(fs["vox"])
        vox.atomic_max = abi.VCT_ACCUM_AVG
        ctx.voxelize(vox)
        alb, nrm = _bytes(ctx.albedo()), _bytes(ctx.normal())
        sc = ob.SceneArrays(app.mesh(), fs["model"])
        ra, rn, st_o = ob.voxelize(sc, fs["lights"], sms, fs["vox"], g.dim, accum=abi.VCT_ACCUM_AVG)
        ra, rn = _bytes(ra), _bytes(rn)
        assert np.array_equal(alb[..., 3], ra[..., 3]) and np.array_equal(nrm[..., 3], rn[..., 3])
        one = ra[..., 3] == 1
        assert one.sum() > 100 and (ra[..., 3] > 1).sum() > 100
        assert np.array_equal(alb[one], ra[one]) and np.array_equal(nrm[one], rn[one])
        # the fragment count does not depend on the order either
        assert ctx.stats().total_voxel_fragments == st_o.total_voxel_fragments
        print(f"{int(one.sum())} single-fragment voxels of {int((ra[..., 3] > 0).sum())}", flush=True)
    finally:
        app.close()


@pytest.mark.parametrize("scene,kw", CASES, ids=["scene1", "scene2_128"])
def test_sum_mode_and_normalize_voxels(scene, kw):
    app, ctx, fs, sms = _setup(scene, kw)
    try:
        g = ctx.grid()
        print(f"scene {scene}: sum voxelization + normalizeVoxels", flush=True)
        vox = abi.VoxelizeParams.from_buffer_copy(fs["vox"])
        vox.atomic_max = abi.VCT_ACCUM_SUM
        ctx.voxelize(vox)
        ctx.normalize_voxels(0.5)
        sc = ob.SceneArrays(app.mesh(), fs["model"])
        sums = np.zeros((g.dim ** 3, 8), np.uint32)
        _, _, st_v = ob.voxelize(sc, fs["lights"], sms, fs["vox"], g.dim, accum=abi.VCT_ACCUM_SUM, sums=sums)
        ra, rn, r0, st = ob.normalize_voxels(sums, g.dim, 0.5)
        assert np.array_equal(ctx.albedo(), ra) and np.array_equal(ctx.normal(), rn)
        assert np.array_equal(ctx.radiance(0).view(np.uint32), ob.half_round(r0).view(np.uint32)
                              if g.storage != abi.VCT_STORE_FP32 else r0.view(np.uint32))
        s = ctx.stats()
        assert (s.unique_voxels, s.max_fragments_per_voxel, s.total_voxel_fragments) == \
            (st.unique_voxels, st.max_fragments_per_voxel, st_v.total_voxel_fragments)
        assert st.max_fragments_per_voxel > 1
        # the usual injection then runs on the normalized grid
        ctx.normalize_inject(fs["inj"])
        assert ctx.stats().lit_voxels > 0
    finally:
        app.close()

"""Level-0 bricks that hold zeros: the injection does not rewrite an all-zero
brick over zeros and the mip chain does not read one; a byte per brick
records which bricks hold zeros, and every other writer of level 0 (upload,
fill holes) resets it.  Consecutive frames of a moving scene, and those
writers between frames, each frame's level 0 and mip chain bit-exact against
the oracle."""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def _inject_frame(app, frame):
    ctx = app.ctx
    fs = app.frame_state(frame)
    ctx.set_transforms(fs["model"])
    sms = []
    for i, L in enumerate(fs["lights"]):
        if L.shadow_caster:
            ctx.shadowmap(i)
            sms.append(ctx.shadow(i))
        else:
            sms.append(None)
    ctx.voxelize(fs["vox"])
    alb = ctx.albedo()
    ctx.normalize_inject(fs["inj"])
    return fs, alb, sms


def _check_level0(app, fs, alb, sms, what, fp32=False):
    ctx = app.ctx
    g = ctx.grid()
    _, ref, _, _ = ob.normalize_inject(alb, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax), fs["lights"], sms,
                                       fp32=fp32)
    got = ctx.radiance(0)
    assert np.array_equal(got.view(np.uint32), ref.reshape(got.shape).view(np.uint32)), \
        f"{what}: {(got != ref.reshape(got.shape)).any(axis=3).sum()} voxels differ"
    return got


def _check_raw_pairs(ctx, rad0, what):
    """Level 0 as stored: fp16 pair elements in 4^3 bricks with a zero border
    brick on each side; element x holds texel x and texel x + 1 (zero past
    the last texel), and the border element before texel 0 holds (0, texel 0).
    The trace's sampler reads both halves, the texel download only the first."""
    d = rad0.shape[0]
    nbp = d // 4 + 2
    raw = ctx.download(abi.VCT_BUF_RADIANCE_RAW, 0, np.float16)
    assert raw.size == nbp ** 3 * 64 * 8
    raw = raw.reshape(nbp, nbp, nbp, 4, 4, 4, 2, 4)  # brick z, y, x; voxel z, y, x; half; channel
    h = rad0.astype(np.float16)  # [z, y, x, c], fp16 values already
    e = np.zeros((d + 8, d + 8, d + 8, 2, 4), np.float16)  # padded texel coordinates, offset 4
    e[4:d + 4, 4:d + 4, 4:d + 4, 0] = h
    e[4:d + 4, 4:d + 4, 4:d + 3, 1] = h[:, :, 1:]
    e[4:d + 4, 4:d + 4, 3, 1] = h[:, :, 0]
    want = e.reshape(nbp, 4, nbp, 4, nbp, 4, 2, 4).transpose(0, 2, 4, 1, 3, 5, 6, 7)
    bad = raw.view(np.uint16) != want.view(np.uint16)
    assert not bad.any(), f"{what}: {int(bad.any(axis=(6, 7)).sum())} stored pair elements differ"


def _zero_bricks(rad0):
    d = rad0.shape[0]
    b = rad0.reshape(d // 4, 4, d // 4, 4, d // 4, 4, 4)
    return ~(b != 0).any(axis=(1, 3, 5, 6))


@pytest.mark.parametrize("storage", [0, 1], ids=["fp16", "fp32"])
def test_moving_scene_over_frames(storage):
    app = make_app(2, dim=128, levels=7, width=480, height=270, shadow_size=256, storage=storage)
    try:
        ctx = app.ctx
        g = ctx.grid()
        prev_zero = None
        changed = 0
        for frame in range(4):
            fs, alb, sms = _inject_frame(app, frame)
            got = _check_level0(app, fs, alb, sms, f"frame {frame}", fp32=(storage == 1))
            if storage == 0:
                _check_raw_pairs(ctx, got, f"frame {frame}")
            ctx.mip(fs["mip"])
            faces = ob.mip_aniso(got, g.levels, fp32=(storage == 1))
            for l in (1, 2):
                for f in range(6):
                    r = ctx.radiance(l, f)
                    assert np.array_equal(r.view(np.uint32), faces[l][f].view(np.uint32)), (frame, l, f)
            zero = _zero_bricks(got)
            if prev_zero is not None:
                assert (zero & prev_zero).any()  # bricks left unwritten
                changed += int((zero != prev_zero).sum())
            prev_zero = zero
        assert changed > 0  # bricks turned empty or occupied between frames
    finally:
        app.close()


def test_upload_and_fill_holes_reset_the_record():
    app = make_app(1, fill_holes=1)
    try:
        ctx = app.ctx
        g = ctx.grid()
        fs, alb, sms = _inject_frame(app, 0)
        got = _check_level0(app, fs, alb, sms, "first injection")
        _check_raw_pairs(ctx, got, "first injection")
        assert _zero_bricks(got).any()
        # an upload writes every brick: the next injection must rewrite the empty ones
        ctx.upload(abi.VCT_BUF_RADIANCE, 0, np.full((g.dim ** 3, 4), 0.25, np.float32))
        fs, alb, sms = _inject_frame(app, 0)
        got = _check_level0(app, fs, alb, sms, "injection after an upload")
        _check_raw_pairs(ctx, got, "injection after an upload")
        # an upload read by the mip chain: no brick is skipped
        up = np.full((g.dim, g.dim, g.dim, 4), 0.5, np.float32)
        ctx.upload(abi.VCT_BUF_RADIANCE, 0, up)
        ctx.mip(fs["mip"])
        lv = ob.mip(ctx.radiance(0), g.levels)
        assert np.array_equal(ctx.radiance(1).view(np.uint32), lv[1].view(np.uint32))
        # fill holes writes into empty bricks: the chain and the next injection see them
        fs, alb, sms = _inject_frame(app, 0)
        rad0 = _check_level0(app, fs, alb, sms, "injection after the mip of an upload")
        ctx.fill_holes()
        filled = ob.fill_holes(rad0)
        assert np.array_equal(ctx.radiance(0).view(np.uint32), filled.view(np.uint32))
        assert (_zero_bricks(rad0) & ~_zero_bricks(filled)).any()  # holes filled in bricks that held zeros
        ctx.mip(fs["mip"])
        lv = ob.mip(filled, g.levels)
        for l in range(1, g.levels):
            assert np.array_equal(ctx.radiance(l).view(np.uint32), lv[l].view(np.uint32)), l
        fs, alb, sms = _inject_frame(app, 0)
        _check_level0(app, fs, alb, sms, "injection after fill holes")
    finally:
        app.close()

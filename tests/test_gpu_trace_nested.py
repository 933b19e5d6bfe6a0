"""The cone tracer skips a finer-level sample where the coarser level's 2x2x2
fetch is zero in every lane, which is exact only in a pyramid built by
injection and filtering (non-negative texels: a filtered texel is zero only
where all the texels it filters are).  An uploaded level 0 may hold negative
values, so after vct_upload the skip must stay off.  Here level 0 holds +v, -v
pairs along x, level 1 filters to exactly zero everywhere, and the traced
image must still match the oracle, which samples every level
(phong.frag:135-180 with LINEAR_MIPMAP_LINEAR)."""
import numpy as np
import pytest

from gpu_util import make_app, rel_err, run_frame_capture
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def test_trace_of_uploaded_signed_level0_samples_every_level():
    app = make_app(1)
    try:
        cap = run_frame_capture(app, 0)
        ctx, g, fs = app.ctx, cap["grid"], cap["fs"]
        assert not g.anisotropic
        d = g.dim
        rng = np.random.default_rng(11)
        half = rng.integers(1, 16, size=(d, d, d // 2, 4)).astype(np.float32) / 64.0  # exact in fp16
        l0 = np.empty((d, d, d, 4), np.float32)
        l0[:, :, 0::2] = half
        l0[:, :, 1::2] = -half
        print(f"upload a signed level 0 ({d}^3), filter, trace", flush=True)
        ctx.upload(abi.VCT_BUF_RADIANCE, 0, l0)
        ctx.mip(fs["mip"])
        levels = [ctx.radiance(l) for l in range(g.levels)]
        assert np.array_equal(levels[0], l0)
        assert not levels[1].any()  # every +v, -v pair cancels
        ctx.trace(fs["trace"])
        got = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
        tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                             app.mesh()["rough"] is not None)
        ref, _ = ob.trace(cap["gbuf"], levels, tp)
        e = rel_err(got[:, :3], ref[:, :3])
        print(f"signed pyramid: indirect mean relative error vs the oracle {e:.3e}", flush=True)
        assert np.abs(ref[:, :3]).sum() > 0  # the finer level contributes: the comparison is not vacuous
        assert e < 1e-3
        assert rel_err(got[:, 3], ref[:, 3]) < 1e-3
        # a frame through injection again enables the skip and still matches
        cap2 = run_frame_capture(app, 0)
        ref2, _ = ob.trace(cap2["gbuf"], cap2["levels"], tp)
        assert rel_err(cap2["indirect"][:, :3], ref2[:, :3]) < 1e-3
    finally:
        app.close()


def test_temporal_injection_keeps_the_skip_off_for_a_signed_level0():
    """The temporal filter keeps decay * the previous value of every unlit
    voxel, so +v, -v pairs uploaded into level 0 survive a temporal injection
    (scaled) and level 1 still cancels to zero over them: the tracer must keep
    sampling the finer level there, as the oracle does."""
    app = make_app(1)
    try:
        cap = run_frame_capture(app, 0)
        ctx, g, fs = app.ctx, cap["grid"], cap["fs"]
        d = g.dim
        rng = np.random.default_rng(12)
        half = rng.integers(1, 16, size=(d, d, d // 2, 4)).astype(np.float32) / 64.0
        l0 = np.empty((d, d, d, 4), np.float32)
        l0[:, :, 0::2] = half
        l0[:, :, 1::2] = -half
        print("upload a signed level 0, temporal injection (decay 0.75), filter, trace", flush=True)
        ctx.upload(abi.VCT_BUF_RADIANCE, 0, l0)
        inj = abi.InjectParams.from_buffer_copy(fs["inj"])
        inj.temporal_filter, inj.temporal_decay = 1, 0.75
        ctx.normalize_inject(inj)
        ctx.mip(fs["mip"])
        levels = [ctx.radiance(l) for l in range(g.levels)]
        assert (levels[0] < 0).any()  # the signed texels survived the temporal mix
        ctx.trace(fs["trace"])
        got = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
        tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                             app.mesh()["rough"] is not None)
        ref, _ = ob.trace(cap["gbuf"], levels, tp)
        e = rel_err(got[:, :3], ref[:, :3])
        print(f"temporal over a signed level 0: indirect mean relative error vs the oracle {e:.3e}", flush=True)
        assert np.abs(ref[:, :3]).sum() > 0
        assert e < 1e-3
        assert rel_err(got[:, 3], ref[:, 3]) < 1e-3
    finally:
        app.close()

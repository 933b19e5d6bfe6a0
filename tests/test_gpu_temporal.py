"""GPU parity of the temporal radiance filter (temporalFilterRadiance,
transferVoxels.comp:55-62 with injectRadiance.comp's overwrite of lit voxels):
level 0 of several consecutive frames, each bit-exact against the oracle fed
the previous frame's level 0 as the GPU stored it."""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene,kw", [(1, {}),
                                      (2, dict(dim=128, levels=7, width=480, height=270, shadow_size=256))],
                         ids=["scene1", "scene2_128"])
def test_temporal_inject_bit_exact_over_frames(scene, kw):
    app = make_app(scene, temporal_filter=1, **kw)
    try:
        ctx = app.ctx
        g = ctx.grid()
        ext = (tuple(g.center), tuple(g.vmin), tuple(g.vmax))
        prev = np.zeros((g.dim ** 3, 4), np.float32)  # the pyramid starts zeroed
        for frame in range(3):
            fs = app.frame_state(frame)
            assert fs["inj"].temporal_filter == 1
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
            got = ctx.radiance(0).reshape(-1, 4)
            _, ref, lm, st = ob.normalize_inject(alb, g.dim, *ext, fs["lights"], sms, prev=prev,
                                                 decay=fs["inj"].temporal_decay)
            ref = ref.reshape(-1, 4)
            print(f"scene {scene} frame {frame}: lit {int(lm.sum())}, "
                  f"carried-over voxels {int(((ref[:, 3] > 0) & (lm == 0)).sum())}", flush=True)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
                f"frame {frame}: {(got != ref).any(axis=1).sum()} voxels differ"
            assert ctx.stats().lit_voxels == st.lit_voxels
            ctx.mip(fs["mip"])  # the chain reads level 0 and leaves it as it is
            prev = got
        if scene == 2:  # moving objects leave decaying radiance behind them
            assert ((ref[:, 3] > 0) & ((alb.reshape(-1) >> 24) == 0)).any()
    finally:
        app.close()

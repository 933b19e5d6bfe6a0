"""GPU parity of the batched shadow-map pass (vct_shadowmaps): every caster's
map bit-identical to the oracle's and to the one-light pass (vct_shadowmap),
across more lights than one rasterizer walk holds."""
import numpy as np
import pytest

from gpu_util import make_app
from oracle import binding as ob
from vct_amd import _abi as abi

pytestmark = pytest.mark.gpu


def _casters(fs):
    return [i for i, L in enumerate(fs["lights"]) if L.shadow_caster]


def test_shadowmaps_match_oracle():
    app = make_app(1)
    try:
        ctx = app.ctx
        fs = app.frame_state(0)
        ctx.set_transforms(fs["model"])
        cast = _casters(fs)
        assert cast
        ctx.shadowmaps()
        sc = ob.SceneArrays(app.mesh(), fs["model"])
        for i in cast:
            got = ctx.shadow(i)
            ref = ob.shadowmap(sc, fs["lights"][i])
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
                f"light {i}: {(got != ref).sum()} texels differ"
            assert (got < 1.0).any()  # the light sees geometry: the comparison is not vacuous
    finally:
        app.close()


def test_shadowmaps_batched_match_one_light_passes():
    """Eleven casters of four map sizes: two walks (8 + 3 lights), and a light
    whose fragments landed in another light's map would show as a mismatch of
    maps whose sizes differ."""
    app = make_app(2)
    try:
        ctx = app.ctx
        fs = app.frame_state(0)
        ctx.set_transforms(fs["model"])
        base = [fs["lights"][i] for i in _casters(fs)]
        assert len(base) >= 2
        sizes = (1024, 512, 768, 256)
        lights = []
        for k in range(11):
            L = abi.Light.from_buffer_copy(base[k % len(base)])
            L.shadow_size = sizes[k % len(sizes)]
            lights.append(L)
        ctx.set_lights(lights)
        ctx.shadowmaps()
        batched = [ctx.shadow(i) for i in range(len(lights))]
        for i in range(len(lights)):
            assert batched[i].shape == (lights[i].shadow_size,) * 2
            ctx.shadowmap(i)
            one = ctx.shadow(i)
            assert np.array_equal(batched[i].view(np.uint32), one.view(np.uint32)), \
                f"light {i} (size {lights[i].shadow_size}): {(batched[i] != one).sum()} texels differ"
            assert (one < 1.0).any()
    finally:
        app.close()


def test_render_uses_every_caster_map():
    """The frame loop's shadow pass leaves each caster's map as the one-light
    pass draws it."""
    app = make_app(2)
    try:
        ctx = app.ctx
        app.render(0)
        fs = app.frame_state(0)
        cast = _casters(fs)
        maps = [ctx.shadow(i) for i in cast]
        for i, m in zip(cast, maps):
            ctx.shadowmap(i)
            assert np.array_equal(m.view(np.uint32), ctx.shadow(i).view(np.uint32)), f"light {i}"
    finally:
        app.close()

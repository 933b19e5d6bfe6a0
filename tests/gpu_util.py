"""Helpers shared by the GPU parity tests: run a frame pass by pass on the GPU,
capturing every intermediate, and rebuild the oracle inputs from them."""
from __future__ import annotations

import numpy as np

import vct_amd
from vct_amd import _abi as abi


def make_app(scene=1, **kw):
    return vct_amd.App(scene, device=0, **kw)


def run_frame_capture(app, frame=0, fill_holes=False):
    """Runs every pass of one frame through the C ABI, downloading the state
    after each pass.  Returns a dict of numpy arrays and the frame state."""
    ctx = app.ctx
    fs = app.frame_state(frame)
    g = ctx.grid()
    out = {"fs": fs, "grid": g}
    ctx.set_transforms(fs["model"])
    lights = fs["lights"]
    sms = []
    for i, L in enumerate(lights):
        if L.shadow_caster:
            ctx.shadowmap(i)
            sms.append(ctx.shadow(i))
        else:
            sms.append(None)
    out["shadow"] = sms
    ctx.voxelize(fs["vox"])
    out["albedo_vox"] = ctx.albedo()
    out["normal_vox"] = ctx.normal()
    out["stats_vox"] = ctx.stats()
    ctx.normalize_inject(fs["inj"])
    out["albedo_norm"] = ctx.albedo()
    out["litmask"] = ctx.litmask()
    out["rad0"] = ctx.radiance(0)
    out["stats"] = ctx.stats()
    if fill_holes:
        ctx.fill_holes()
        out["rad0_filled"] = ctx.radiance(0)
    ctx.mip(fs["mip"])
    out["levels"] = [ctx.radiance(l) for l in range(g.levels)]
    if fs["mip"].albedo_chain:
        out["albedo_levels"] = [None] + [
            ctx.download(abi.VCT_BUF_ALBEDO, l, np.uint32).reshape((max(1, g.dim >> l),) * 3)
            for l in range(1, g.levels)]
    if g.anisotropic:
        out["faces"] = [None] + [[ctx.radiance(l, f) for f in range(6)] for l in range(1, g.levels)]
    ctx.gbuffer(fs["cam"])
    cam = fs["cam"]
    out["gbuf"] = ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(cam.height, cam.width, 16)
    ctx.trace(fs["trace"])
    out["indirect"] = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)
    if fs["trace"].enable_reflections:
        out["specular"] = ctx.download(abi.VCT_BUF_SPECULAR, 0, np.float32).reshape(-1, 4)
    cp = app.compose_params(frame)
    ctx.compose(cp)
    out["compose_params"] = cp
    out["color"] = ctx.color()
    return out


def compose_variant(base: abi.ComposeParams, **kw) -> abi.ComposeParams:
    p = abi.ComposeParams.from_buffer_copy(base)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def compose_reference(app, cap, params):
    """The oracle's composition on the G-buffer, trace output and shadow maps
    the GPU produced (so only the compose pass itself is compared)."""
    from oracle import binding as ob
    fs = cap["fs"]
    return ob.compose(cap["gbuf"], cap["indirect"], cap.get("specular"), fs["lights"], cap["shadow"], params,
                      has_roughness=app.mesh()["rough"] is not None, occluded=bool(fs["trace"].draw_occlusion))


def assert_compose_close(got, ref, what):
    # identical operation order; powf and sqrtf may round differently by an ulp
    # between the GPU math library and the C library, so the bar is 1e-5 relative
    bad = ~np.isclose(got, ref, rtol=1e-5, atol=1e-6)
    assert not bad.any(), f"{what}: {int(bad.sum())} channels differ, max abs {np.abs(got - ref).max():.3e}"


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """Mean relative error sum|a-b| / sum|b| (the image-level tolerance metric)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())

"""Host side of the learned diffuse gather: the training-point recipe and the
held-out scoring used by the benchmark and the tests.

No reference counterpart (the reference has no learned stage).  Everything
here drives the C ABI through vct_amd.App / Context: the labels and the
64-cone reference come from vct_learned_prepare / vct_reference_gather on the
GPU, the network from vct_learned_train.
"""
from __future__ import annotations

import numpy as np

from . import _abi as abi

TRAIN_SEED0 = 100   # training cameras: seeds 100, 101, ...
VALIDATION_SEEDS = tuple(range(200, 208))
TEST_SEEDS = tuple(range(300, 316))


def training_points(app, n: int = 1 << 20, view_fraction: float = 0.75, train_cameras: int = 32,
                    area_seed: int = 7, rng_seed: int = 3) -> np.ndarray:
    """n G-buffer texels to train on: round(n * view_fraction) drawn (seeded,
    with replacement only if there are too few) from the covered texels of
    the 256x256 G-buffers of `train_cameras` seeded cameras, the rest
    area-uniform surface points of frame 0; shuffled.  Leaves the context's
    G-buffer at the last training camera."""
    ctx = app.ctx
    rng = np.random.default_rng(rng_seed)
    nv = int(n * view_fraction)
    area = app.surface_points(n - nv, 0, seed=area_seed) if n > nv else np.zeros((0, 16), np.float32)
    if nv:
        views = []
        for sd in range(TRAIN_SEED0, TRAIN_SEED0 + train_cameras):
            ctx.gbuffer(app.camera(sd, 256, 256))
            g = ctx.download(abi.VCT_BUF_GBUFFER, 0, np.float32).reshape(-1, 16)
            views.append(g[g[:, 11].view(np.uint32) == 1])
        view = np.concatenate(views)
        pts = np.concatenate([area, view[rng.choice(len(view), nv, replace=len(view) < nv)]])
    else:
        pts = area
    return np.ascontiguousarray(pts[rng.permutation(len(pts))])


def held_out_views(app, seeds, width: int = 480, height: int = 270, scene_camera: bool = False):
    """(name, camera) pairs: optionally the scene's own camera, then the
    seeded cameras `seeds`, all at width x height."""
    out = []
    if scene_camera:
        c = abi.Camera.from_buffer_copy(app.frame_state(0)["cam"])
        c.width, c.height = width, height
        out.append(("scene", c))
    return out + [(f"seed{sd}", app.camera(sd, width, height)) for sd in seeds]


def pooled_error(app, trace_params, views, threshold: float = 1e30, plain: bool = False, refs=None) -> dict:
    """Pooled mean relative error sum|out - ref| / sum|ref| over the rgb of
    every pixel of every view, against the 64-cone reference gather, of the
    learned gather (tiles whose predicted-luminance variance exceeds
    `threshold` fall back to the 6-cone tracer) or, with plain=True, of the
    6-cone tracer alone.  `refs` (a dict) caches the references by view name."""
    ctx = app.ctx
    refs = {} if refs is None else refs
    num = den = 0.0
    per, nfb, ntiles = {}, 0, 0
    for name, cam in views:
        ctx.gbuffer(cam)
        if name not in refs:
            ctx.reference_gather()
            refs[name] = ctx.download(abi.VCT_BUF_REFERENCE, 0, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
        ref = refs[name]
        if plain:
            ctx.trace(trace_params)
        else:
            nfb += ctx.trace_learned(trace_params, threshold)
        ntiles += ((cam.width + 7) // 8) * ((cam.height + 7) // 8)
        out = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
        a, b = float(np.abs(out - ref).sum()), float(np.abs(ref).sum())
        num += a
        den += b
        per[name] = {"err": round(a / b, 4) if b > 0 else None, "mean_ref": round(b / (3 * ref.shape[0]), 6)}
    return {"pooled": round(num / den, 5) if den > 0 else None, "fallback_tile_fraction": round(nfb / ntiles, 4),
            "per_view": per}


def in_sample_error(app, trace_params, points: np.ndarray, n: int = 65536) -> float:
    """The pooled error of the learned gather on the first n training points
    themselves, uploaded as a square G-buffer."""
    ctx = app.ctx
    side = int(np.sqrt(n))
    g = np.ascontiguousarray(points[: side * side])
    cam = abi.Camera.from_buffer_copy(app.frame_state(0)["cam"])
    cam.width = cam.height = side
    ctx.gbuffer(cam)
    ctx.upload(abi.VCT_BUF_GBUFFER, 0, g)
    ctx.reference_gather()
    ref = ctx.download(abi.VCT_BUF_REFERENCE, 0, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
    ctx.trace_learned(trace_params, 1e30)
    out = ctx.download(abi.VCT_BUF_INDIRECT, 0, np.float32).reshape(-1, 4)[:, :3].astype(np.float64)
    return round(float(np.abs(out - ref).sum() / np.abs(ref).sum()), 5)


def calibrate_threshold(app, trace_params, fraction: float, seed: int = TRAIN_SEED0, width: int = 1920,
                        height: int = 1080) -> tuple[float, float]:
    """The variance threshold at which `fraction` of the tiles of training
    camera `seed` (width x height) fall back, by bisection on log10 of the
    threshold; returns (threshold, achieved fraction)."""
    ctx = app.ctx
    ctx.gbuffer(app.camera(seed, width, height))
    tiles = ((width + 7) // 8) * ((height + 7) // 8)
    lo, hi = -12.0, 2.0
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        if ctx.trace_learned(trace_params, 10.0 ** mid) > fraction * tiles:
            lo = mid
        else:
            hi = mid
    return 10.0 ** hi, ctx.trace_learned(trace_params, 10.0 ** hi) / tiles

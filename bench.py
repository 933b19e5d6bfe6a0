"""Benchmark: per-frame voxelize + inject + mip + cone trace on one MI355X.

Workload: the full-size dynamic scene (2000 boxes + 64 subdivision-4
icospheres = 351,680 triangles moving rigidly every frame, 256^3 grid with 8
levels of 6-direction fp16 mips, 1 directional + 4 point lights with 1024^2
shadow maps, 6 diffuse + 1 specular cone at 1920x1080).  A step is one frame:
every object transform changes, and the scene is re-voxelized, re-injected,
re-filtered and re-traced.

Prints one JSON line.  `value` is ms per frame of the four metric passes
(voxelize + normalize/inject + mip + trace), from HIP events on the context
stream; the shadow-map and G-buffer passes the frame also runs are reported
beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scene", type=int, default=2)
    ap.add_argument("--storage", type=int, default=-1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        return 2

    import vct_amd

    app = vct_amd.App(args.scene, device=0, storage=args.storage)
    ctx = app.ctx
    cfg = app.cfg
    ntri, nobj, nlights = app.sizes()
    for f in range(args.warmup):
        app.render(f)
    ctx.sync()
    ctx.pass_times()  # discard warm-up timings
    t0 = time.perf_counter()
    for f in range(args.steps):
        app.render(args.warmup + f)
    ctx.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    pt = ctx.pass_times()
    k = float(args.steps)
    passes = {
        "shadowmap_ms": pt.shadowmap_ms / k,
        "voxelize_ms": pt.voxelize_ms / k,
        "radiance_ms": pt.radiance_ms / k,
        "mipmap_ms": pt.mipmap_ms / k,
        "gbuffer_ms": pt.gbuffer_ms / k,
        "render_ms": pt.render_ms / k,
    }
    metric_ms = passes["voxelize_ms"] + passes["radiance_ms"] + passes["mipmap_ms"] + passes["render_ms"]
    npix = cfg.width * cfg.height
    cones = npix * (6 + (1 if cfg.reflections else 0))
    out = {
        "metric": "ms/frame voxelize+inject+mip+trace @256^3,1080p; GB/s vs peak; learned vs plain",
        "value": round(metric_ms, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 accumulate, fp16 storage" if cfg.storage == 0 else ("fp32" if cfg.storage == 1 else "fp8"),
        "data": "synthetic",
        "config": {"workload": f"scene {args.scene}: {ntri} tris, {nobj} objects, {nlights} lights, "
                               f"{cfg.dim}^3 x {cfg.levels} levels, {'6-dir' if cfg.anisotropic else 'iso'} mips, "
                               f"{cfg.width}x{cfg.height}, {6 + (1 if cfg.reflections else 0)} cones/px",
                   "parallelism": "single GPU"},
        "passes": {k2: round(v, 4) for k2, v in passes.items()},
        "frame_ms_all_passes": round(sum(passes.values()), 4),
        "cones_per_s": round(cones / (passes["render_ms"] * 1e-3), 1) if passes["render_ms"] > 0 else None,
    }
    print(json.dumps(out), flush=True)
    app.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

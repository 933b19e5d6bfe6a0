"""Writes the golden fixtures of the seeded scenes and the oracle's scene-1 passes.

    python tests/golden/make_golden.py

scene1.npz   the small scene's geometry (36 triangles), frame-0 object
             matrices, light-space matrix, camera and every oracle pass output
             digest (sha256 of the raw bytes) at 64^3.
scene2.json  sha256 digests of the full-size scene's geometry and frame-0
             matrices (the arrays themselves are 25 MB).

The reference ships no tests or vectors; these fixtures pin this build's own
scene generator and oracle against unintended change.  Regenerate them (and
say why in the commit) whenever the scene definition or an oracle pass
changes on purpose.
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import vct_amd  # noqa: E402
from oracle import binding as ob  # noqa: E402

HERE = Path(__file__).resolve().parent


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene1_passes(app):
    """The oracle's pass outputs for frame 0 of the small scene: exact bytes of
    the integer and storage-rounded passes, float sums of the G-buffer and the
    traced images."""
    cfg = app.cfg
    mesh, fs = app.mesh(), app.frame_state(0)
    sc = ob.SceneArrays(mesh, fs["model"])
    L = fs["lights"]
    sms = [ob.shadowmap(sc, l) if l.shadow_caster else None for l in L]
    alb, nrm, st = ob.voxelize(sc, L, sms, fs["vox"], cfg.dim)
    h = 1.0 + 1.0 / (cfg.dim - 1)  # the small scene's grid extent
    alb2, rad0, lm, st2 = ob.normalize_inject(alb, cfg.dim, (0, 0, 0), (-h,) * 3, (h,) * 3, L, sms)
    levels = ob.mip(rad0, cfg.levels)
    gb = ob.gbuffer(sc, fs["cam"])
    tp = ob.trace_params(fs["trace"], cfg.dim, cfg.levels, (0, 0, 0), (-h,) * 3, (h,) * 3, False)
    ind, spec = ob.trace(gb, levels, tp)
    g = gb.reshape(-1, 16)
    bits = [3, 11, 15]  # albedo, covered flag, triangle id: integer bit patterns
    floats = [c for c in range(16) if c not in bits]
    exact = dict(shadow0=sms[0], albedo_vox=alb, normal_vox=nrm, albedo=alb2, litmask=lm, radiance0=rad0,
                 gbuffer_bits=np.ascontiguousarray(g[:, bits]).view(np.uint32),
                 mips=np.concatenate([l.reshape(-1) for l in levels[1:]]),
                 stats=np.asarray([st.total_voxel_fragments, st2.unique_voxels, st2.max_fragments_per_voxel,
                                   st2.lit_voxels], np.uint32))
    sums = dict(gbuffer=g[:, floats].astype(np.float64).sum(0),
                indirect=ind.astype(np.float64).sum(0), specular=spec.astype(np.float64).sum(0))
    return exact, sums


def main():
    app = vct_amd.App(1, device=-1)
    m, fs = app.mesh(), app.frame_state(0)
    out = dict(pos=m["pos"], nrm=m["nrm"], tan=m["tan"], albedo=m["albedo"], obj=m["obj"], model=fs["model"],
               light_ls=np.asarray(list(fs["lights"][0].ls), np.float32),
               cam_view=np.asarray(list(fs["cam"].view), np.float32),
               cam_proj=np.asarray(list(fs["cam"].projection), np.float32))
    exact, sums = scene1_passes(app)
    for k, v in exact.items():
        out["digest_" + k] = np.frombuffer(bytes.fromhex(digest(v)), np.uint8)
    for k, v in sums.items():
        out["sum_" + k] = v
    np.savez(HERE / "scene1.npz", **out)
    app2 = vct_amd.App(2, device=-1)
    m2, fs2 = app2.mesh(), app2.frame_state(0)
    j = {k: digest(m2[k]) for k in ("pos", "nrm", "tan", "albedo", "obj", "rough")}
    j["model_frame0"] = digest(fs2["model"])
    j["model_frame7"] = digest(app2.frame_state(7)["model"])
    j["ntri"], j["nobj"], j["nlights"] = app2.sizes()
    (HERE / "scene2.json").write_text(json.dumps(j, indent=1) + "\n")
    print("wrote", HERE / "scene1.npz", HERE / "scene2.json")


if __name__ == "__main__":
    main()

"""A scene loaded from an OBJ/MTL file through the GPU passes: shadow map,
voxelization and injection bit-exact against the oracle on the same loaded
mesh, the trace within the indirect-image tolerance, and a composed frame."""
import numpy as np
import pytest

import vct_amd
from gpu_util import rel_err, run_frame_capture
from obj_fixtures import write_cube
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def test_obj_scene_passes_match_oracle(tmp_path):
    app = vct_amd.App(obj_path=write_cube(tmp_path), device=0, width=320, height=240)
    try:
        print("obj scene: frame capture", flush=True)
        cap = run_frame_capture(app, 0)
        fs, g = cap["fs"], cap["grid"]
        mesh = app.mesh()
        sc = ob.SceneArrays(mesh, fs["model"])
        ref_sm = ob.shadowmap(sc, fs["lights"][0])
        assert np.array_equal(cap["shadow"][0].view(np.uint32), ref_sm.view(np.uint32))
        alb, nrm, st = ob.voxelize(sc, fs["lights"], cap["shadow"], fs["vox"], g.dim)
        assert np.array_equal(cap["albedo_vox"], alb) and np.array_equal(cap["normal_vox"], nrm)
        assert cap["stats_vox"].total_voxel_fragments == st.total_voxel_fragments > 0
        _, rad0, lm, _ = ob.normalize_inject(alb, g.dim, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                                             fs["lights"], cap["shadow"])
        assert np.array_equal(cap["rad0"].view(np.uint32), rad0.view(np.uint32))
        tp = ob.trace_params(fs["trace"], g.dim, g.levels, tuple(g.center), tuple(g.vmin), tuple(g.vmax),
                             mesh["rough"] is not None)
        ind, _ = ob.trace(cap["gbuf"], cap["levels"], tp)
        e = rel_err(cap["indirect"][:, :3], ind[:, :3])
        print(f"obj scene: indirect mean relative error vs the FP32 oracle {e:.3e}", flush=True)
        assert e < 1e-3 and ind[:, :3].sum() > 0
        col = cap["color"]
        cov = cap["gbuf"][..., 11].view(np.uint32).reshape(-1) != 0
        assert cov.mean() > 0.05 and np.isfinite(col).all() and col[cov, :3].sum() > 0
    finally:
        app.close()

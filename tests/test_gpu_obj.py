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


def test_obj_vertex_tangents_reach_the_gbuffer(tmp_path):
    """A smooth UV sphere whose vertices are shared by their faces: the
    G-buffer interpolates the per-vertex tangent sums (Mesh.cpp:190-197,
    phong.vert:48-52), bit-exact against the oracle given the same tangents,
    and differs from the image the per-face tangents would give."""
    from obj_fixtures import uv_sphere_obj
    p = tmp_path / "sphere.obj"
    p.write_text(uv_sphere_obj(16, 32))
    app = vct_amd.App(obj_path=p, device=0, width=256, height=192)
    try:
        print("obj sphere: G-buffer with vertex tangents", flush=True)
        cap = run_frame_capture(app, 0)
        fs = cap["fs"]
        mesh = app.mesh()
        assert mesh["vtan"] is not None
        ref = ob.gbuffer(ob.SceneArrays(mesh, fs["model"]), fs["cam"])
        got = cap["gbuf"]
        cov = got[..., 11].view(np.uint32) != 0
        assert cov.mean() > 0.1
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
            f"attributes max abs diff {np.abs(got - ref).max()}"
        flat = dict(mesh, vtan=None)
        ref_face = ob.gbuffer(ob.SceneArrays(flat, fs["model"]), fs["cam"])
        assert np.abs(ref_face[cov][:, 8:11] - got[cov][:, 8:11]).max() > 1e-3  # the tangent columns
    finally:
        app.close()

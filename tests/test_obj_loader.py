"""OBJ/MTL ingestion (vct_obj_load; Mesh::loadMesh, src/Graphics/Mesh.cpp:42-266)
and the application built on a loaded file, on the CPU."""
import numpy as np
import pytest

import vct_amd
from obj_fixtures import PENTAGON_OBJ, write_cube


def _rgba8(c):
    return int(round(c[0] * 255)) | (int(round(c[1] * 255)) << 8) | (int(round(c[2] * 255)) << 16) | (255 << 24)


def test_cube_faces_materials_normals(tmp_path):
    m = vct_amd.load_obj(write_cube(tmp_path))
    assert m["pos"].shape == (12, 3, 3) and m["materials"] == 2
    assert np.array_equal(m["bmin"], [-1, -1, -1]) and np.array_equal(m["bmax"], [1, 1, 1])
    assert list(m["albedo"][:4]) == [_rgba8((0.8, 0.1, 0.2))] * 4      # +z and -z faces: red
    assert list(m["albedo"][4:]) == [_rgba8((0.5, 0.5, 0.5))] * 8      # the rest: grey
    # every vertex normal is the face's given normal, and points away from the centre
    centroid = m["pos"].mean(axis=1)
    for t in range(12):
        assert np.allclose(m["nrm"][t], m["nrm"][t][0])
        assert np.dot(m["nrm"][t][0], centroid[t]) > 0
    # the last face uses negative (relative) indices: the -y face
    assert np.allclose(m["pos"][10:].reshape(-1, 3)[:, 1], -1)
    assert m["rough"] is not None and np.isclose(m["rough"][0], 0.3)


def test_cube_tangent_follows_the_uv_rule(tmp_path):
    # +z face: u runs along +x and (flipped) v along -y, so T = +x (Mesh.cpp:174-189)
    m = vct_amd.load_obj(write_cube(tmp_path))
    assert np.allclose(m["tan"][0], [1, 0, 0], atol=1e-6) and np.allclose(m["tan"][1], [1, 0, 0], atol=1e-6)
    t = m["tan"]
    assert np.allclose(np.linalg.norm(t, axis=1), 1, atol=1e-6)
    assert np.allclose(np.einsum("ij,ij->i", t, m["nrm"][:, 0]), 0, atol=1e-6)  # tangent-plane vectors


def test_polygon_fan_and_missing_attributes(tmp_path):
    p = tmp_path / "pent.obj"
    p.write_text(PENTAGON_OBJ)
    m = vct_amd.load_obj(p)
    assert m["pos"].shape[0] == 3                                       # fan (0,1,2) (0,2,3) (0,3,4)
    assert np.array_equal(m["pos"][:, 0], np.zeros((3, 3), np.float32))
    assert np.allclose(m["nrm"], [0, 0, 1])                             # face normal
    assert np.allclose(np.linalg.norm(m["tan"], axis=1), 1) and np.allclose(m["tan"][:, 2], 0)
    assert np.all(m["albedo"] == 0xffffffff) and m["rough"] is None    # no material: white


@pytest.mark.parametrize("text,needle", [("v 0 0 0\nv 1 0 0\nf 1 2 3\n", "bad face at line 3"),
                                         ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n", "bad face"),
                                         ("v 0 0 0\n", "no faces"),
                                         ("mtllib missing.mtl\n", "cannot open material library")])
def test_faulty_files_are_reported(tmp_path, text, needle):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(vct_amd.VCTError) as e:
        vct_amd.load_obj(p)
    assert needle in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(vct_amd.VCTError):
        vct_amd.load_obj(tmp_path / "none.obj")


def test_app_frames_the_loaded_mesh(oracle, tmp_path):
    app = vct_amd.App(obj_path=write_cube(tmp_path), device=-1)
    fs = app.frame_state(0)
    half = np.float32(1.05)
    expect = oracle.voxel_projections((0, 0, 0), (-half,) * 3, (half,) * 3)
    assert np.array_equal(np.asarray(list(fs["vox"].mvp_x), np.float32), expect[0])
    assert np.array_equal(fs["model"][0], np.eye(4, dtype=np.float32).reshape(-1))  # one static actor
    eye = np.asarray(list(fs["cam"].eye))
    assert np.max(np.abs(eye)) > 1                                     # the camera is outside the cube
    # the loaded mesh voxelizes to a closed shell inside the grid
    sc = oracle.SceneArrays(app.mesh(), fs["model"])
    alb, _, st = oracle.voxelize(sc, fs["lights"], [None], fs["vox"], 64, lighting=0)
    occ = (alb >> 24) > 0
    assert st.total_voxel_fragments > 0
    assert np.array_equal(occ, occ[::-1, :, :]) and np.array_equal(occ, occ[:, :, ::-1])  # symmetric cube
    assert not occ[32, 32, 32]                                          # hollow


def test_sphere_vertex_tangents_are_face_sums(tmp_path):
    """Mesh.cpp:138-197: a vertex shared by several faces (same v/vt/vn
    triple) gets the normalized sum of their tangents.  On a UV sphere whose u
    runs with the longitude, that is the east direction (-sin phi, 0, cos phi)
    up to the faces' curvature, the same for every corner of the vertex, and
    different from its faces' own tangents."""
    import math
    from obj_fixtures import uv_sphere_obj
    p = tmp_path / "sphere.obj"
    p.write_text(uv_sphere_obj(12, 24))
    m = vct_amd.load_obj(p)
    vt, ft, pos = m["vtan"], m["tan"], m["pos"].astype(np.float64)
    assert vt.shape == (m["pos"].shape[0], 3, 3)
    assert np.allclose(np.linalg.norm(vt, axis=2), 1, atol=1e-6)
    # the corners of one vertex (same position, away from the seam) share its tangent
    keys = {}
    for t in range(pos.shape[0]):
        for c in range(3):
            k = tuple(np.round(pos[t, c], 5))
            keys.setdefault(k, []).append(vt[t, c])
    shared = [np.asarray(v) for k, v in keys.items() if len(v) > 3 and abs(k[1]) < 0.95]
    assert len(shared) > 100
    # the seam column (u = 0 and u = 1) splits vertices: compare away from x = 1, z = 0
    inner = [(k, v) for k, v in keys.items() if abs(k[1]) < 0.95 and not (k[0] > 0.2 and abs(k[2]) < 1e-4)]
    for k, v in inner:
        v = np.asarray(v)
        assert np.allclose(v, v[0], atol=1e-6), k
        phi = math.atan2(k[2], k[0])
        east = np.array([-math.sin(phi), 0.0, math.cos(phi)])
        assert np.dot(v[0], east) > 0.995, (k, v[0], east)
    # per-face tangents of the faces around a vertex differ from its sum
    assert np.abs(vt - ft[:, None, :]).max() > 0.05

"""The seeded scenes and the oracle's scene-1 passes against the committed
fixtures (tests/golden/make_golden.py writes them)."""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import vct_amd

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402


def _digest(a) -> bytes:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest()


@pytest.fixture(scope="module")
def g1():
    return np.load(GOLDEN / "scene1.npz")


@pytest.fixture(scope="module")
def app1():
    return vct_amd.App(1, device=-1)


def test_scene1_geometry_and_frame_state(g1, app1):
    m, fs = app1.mesh(), app1.frame_state(0)
    for k in ("pos", "nrm", "tan", "albedo", "obj"):
        assert np.array_equal(m[k], g1[k]), k
    assert np.array_equal(fs["model"], g1["model"])
    assert np.array_equal(np.asarray(list(fs["lights"][0].ls), np.float32), g1["light_ls"])
    assert np.array_equal(np.asarray(list(fs["cam"].view), np.float32), g1["cam_view"])
    assert np.array_equal(np.asarray(list(fs["cam"].projection), np.float32), g1["cam_proj"])


def test_scene1_oracle_passes(g1, app1):
    exact, sums = make_golden.scene1_passes(app1)
    for k, v in exact.items():
        assert _digest(v) == g1["digest_" + k].tobytes(), k
    for k, v in sums.items():
        assert np.allclose(v, g1["sum_" + k], rtol=1e-9, atol=0), k


def test_scene2_digests():
    j = json.loads((GOLDEN / "scene2.json").read_text())
    app = vct_amd.App(2, device=-1)
    m = app.mesh()
    assert list(app.sizes()) == [j["ntri"], j["nobj"], j["nlights"]]
    for k in ("pos", "nrm", "tan", "albedo", "obj", "rough"):
        assert _digest(m[k]).hex() == j[k], k
    assert _digest(app.frame_state(0)["model"]).hex() == j["model_frame0"]
    assert _digest(app.frame_state(7)["model"]).hex() == j["model_frame7"]

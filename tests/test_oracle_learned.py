"""Pins the learned-gather references: the FP64 network's gradients against
central finite differences, BF16 rounding, Adam against its closed form on a
scalar, and the 64-cone label integral (or_label.c) against closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import mlp_ref


def test_fp64_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    params = mlp_ref.init_params(1)
    # move biases off zero so every ReLU path is exercised
    params = [(W, rng.normal(0, 0.1, b.shape)) for W, b in params]
    x = rng.normal(size=(8, 129))
    t = rng.normal(size=(8, 3))
    _, grads = mlp_ref.loss_and_grads(params, x, t)
    flat = mlp_ref.flatten(params)
    g = mlp_ref.flatten([(gW, gb) for gW, gb in grads])
    idx = rng.choice(flat.size, 200, replace=False)
    eps = 1e-6
    for i in idx:
        p = flat.copy()
        p[i] += eps
        lp, _ = mlp_ref.loss_and_grads(mlp_ref.unflatten(p), x, t)
        p[i] -= 2 * eps
        lm, _ = mlp_ref.loss_and_grads(mlp_ref.unflatten(p), x, t)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_param_count():
    assert mlp_ref.NPARAMS == 16835
    assert mlp_ref.flatten(mlp_ref.init_params(0)).size == 16835


def test_bf16_round_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-39, 65504.0], np.float32)
    out = mlp_ref.bf16_round(vals)
    # 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7: ties to even -> 1.0
    assert out[1] == 1.0
    assert out[2] == 1.0078125  # above the midpoint
    assert out[0] == 1.0 and out[3] == 1.0078125 and out[4] == -2.5
    for v in np.random.default_rng(1).normal(size=1000).astype(np.float32):
        r = mlp_ref.bf16_round(np.array([v]))[0]
        assert abs(r - v) <= abs(v) * 2 ** -8 + 1e-40
        assert (np.float32(r).view(np.uint32) & 0xFFFF) == 0


def test_adam_first_steps_closed_form():
    params = [(np.array([[1.0]]), np.array([0.0]))] * 4
    grads = [(np.array([[0.5]]), np.array([0.0]))] * 4
    st = mlp_ref.adam_state(params)
    p1 = mlp_ref.adam_step(params, grads, st, 1, lr=0.1)
    # step 1: m_hat = g, v_hat = g^2 -> update lr * g / (|g| + eps)
    assert abs(p1[0][0][0, 0] - (1.0 - 0.1 * 0.5 / (0.5 + 1e-8))) < 1e-12


def _label_lib(oracle):
    L = oracle.lib()
    L.or_label64.restype = None
    L.or_label64.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p]
    L.or_label_dirs.restype = None
    L.or_label_dirs.argtypes = [C.c_void_p, C.c_void_p]
    return L


def test_label_directions_and_weights(oracle):
    L = _label_lib(oracle)
    d = np.zeros((64, 3))
    w = np.zeros(64)
    L.or_label_dirs(d.ctypes.data, w.ctypes.data)
    assert np.allclose(np.linalg.norm(d, axis=1), 1)
    assert np.all(d[:, 1] > 0)  # the hemisphere around local +y
    assert abs(w.sum() - 1) < 1e-12
    # uniform in cos(theta): cosine-weighted mean of y is 2/3 for the hemisphere
    assert abs((w * d[:, 1]).sum() - 2 / 3) < 1e-3


def _label(oracle, levels, g):
    L = _label_lib(oracle)
    lp = (C.c_void_p * len(levels))(*[a.ctypes.data for a in levels])
    out = np.zeros((g.shape[0], 4))
    c = np.zeros(3, np.float32)
    mn = -np.ones(3, np.float32)
    mx = np.ones(3, np.float32)
    L.or_label64(g.ctypes.data, g.shape[0], C.cast(lp, C.c_void_p), levels[0].shape[0], len(levels), c.ctypes.data,
                 mn.ctypes.data, mx.ctypes.data, out.ctypes.data)
    return out


def _point():
    g = np.zeros((1, 16), np.float32)
    g[0, 4:7] = [0, 0, 1]      # normal
    g[0, 8:11] = [1, 0, 0]     # tangent
    g[0, 12:15] = [0, 1, 0]    # bitangent (the gather axis)
    g[0, 11] = np.frombuffer(np.uint32(1).tobytes(), np.float32)[0]
    return g


def test_label_of_empty_volume_is_zero(oracle):
    levels = [np.zeros((16 >> l,) * 3 + (4,), np.float32) for l in range(5)]
    out = _label(oracle, levels, _point())
    assert np.all(out[0, :3] == 0) and out[0, 3] == 1


def test_label_of_uniform_opaque_volume(oracle):
    """Every level uniformly (c, a=1): the first sample of every cone is fully
    opaque after the half-step correction 1 - (1 - 1)^(1/2) = 1, so each cone
    returns (c/2, 1) and the gather returns (c/2) * occlusion 0 = 0, with
    occlusion 0; with a = 0.75 the first sample gives alpha 0.5."""
    c = 0.4
    for a, alpha1 in ((1.0, 1.0), (0.75, 0.5)):
        levels = []
        for l in range(5):
            v = np.zeros((16 >> l,) * 3 + (4,), np.float32)
            v[...] = [c, c, c, a]
            levels.append(v)
        out = _label(oracle, levels, _point())
        if a == 1.0:
            assert abs(out[0, 3]) < 1e-9 and np.all(np.abs(out[0, :3]) < 1e-9)
        else:
            assert 0 <= out[0, 3] < 0.5  # several samples before exit or border

"""FP64 numpy reference of the learned diffuse gather's network.

TEST INFRASTRUCTURE ONLY (the checker of the HIP MLP kernels; the product
never imports it).

Network: 129 -> 64 -> 64 -> 64 -> 3, ReLU on the hidden layers, weights
W_k of shape (out, in) and biases b_k.  Loss: mean squared error over the
batch and the 3 outputs.  Optimizer: Adam (beta1 0.9, beta2 0.999, eps 1e-8)
on FP32 master weights.

With `bf16=True` the reference rounds (round to nearest even) at exactly the
points where the GPU kernels feed MFMA: every matmul operand -- layer inputs,
weights, and the back-propagated gradients -- becomes bfloat16, while every
sum stays in FP64 (the GPU accumulates in FP32).  With `bf16=False` it is a
plain FP64 network, checked against central finite differences.
"""
from __future__ import annotations

import numpy as np

SIZES = [129, 64, 64, 64, 3]


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to nearest even bfloat16, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(f)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def init_params(seed: int = 0):
    """He-uniform weights, zero biases (the GPU trainer's initialisation)."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(4):
        fan_in, fan_out = SIZES[i], SIZES[i + 1]
        lim = np.sqrt(6.0 / fan_in)
        params.append((rng.uniform(-lim, lim, (fan_out, fan_in)), np.zeros(fan_out)))
    return params


def forward(params, x, bf16=False):
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    acts = [np.asarray(x, np.float64)]
    h = acts[0]
    for i, (W, b) in enumerate(params):
        z = r(h) @ r(W).T + b
        h = np.maximum(z, 0.0) if i < 3 else z
        acts.append(h)
    return h, acts


def loss_and_grads(params, x, t, bf16=False):
    r = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    y, acts = forward(params, x, bf16)
    n = y.size
    loss = float(((y - t) ** 2).sum() / n)
    dz = 2.0 * (y - t) / n
    grads = [None] * 4
    for i in range(3, -1, -1):
        W, _ = params[i]
        a = acts[i]
        grads[i] = (r(dz).T @ r(a), dz.sum(0))
        if i > 0:
            da = r(dz) @ r(W)
            dz = da * (acts[i] > 0)
    return loss, grads


def adam_step(params, grads, state, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; `state` is a list of (m, v) pairs per tensor (in/out)."""
    new = []
    k = 0
    for (W, b), (gW, gb) in zip(params, grads):
        out = []
        for p, g in ((W, gW), (b, gb)):
            m, v = state[k]
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            mh = m / (1 - b1 ** step)
            vh = v / (1 - b2 ** step)
            out.append(p - lr * mh / (np.sqrt(vh) + eps))
            state[k] = (m, v)
            k += 1
        new.append(tuple(out))
    return new


def adam_state(params):
    st = []
    for W, b in params:
        st.append((np.zeros_like(W), np.zeros_like(W)))
        st.append((np.zeros_like(b), np.zeros_like(b)))
    return st


def flatten(params) -> np.ndarray:
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params])


def unflatten(v: np.ndarray):
    out, o = [], 0
    for i in range(4):
        fi, fo = SIZES[i], SIZES[i + 1]
        W = v[o:o + fo * fi].reshape(fo, fi)
        o += fo * fi
        b = v[o:o + fo]
        o += fo
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


NPARAMS = sum(SIZES[i] * SIZES[i + 1] + SIZES[i + 1] for i in range(4))

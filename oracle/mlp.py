"""ORACLE — test infrastructure only.

NumPy restatement of the learned scheduler's network (dynamicgraphs_scc_amd/
csrc/device/sched.hip): 16 -> 64 -> 64 -> 2, ReLU, heads out0 (regression) and
out1 (logit).  Loss per sample: M0 (out0 - Y0)^2 + M1 BCE(sigmoid(out1), Y1),
averaged over the samples.  With emulate_bf16 the operands of every matrix
product are rounded to bfloat16 (round to nearest even) exactly where the GPU
kernels round them; sums are in float64 here and FP32 on the GPU.

The network has no counterpart in the reference.  Pinning (tests/test_sched.py):
the backward pass equals central finite differences of the forward pass
(float64, no rounding); the forward pass equals hand-computed outputs of
hand-set weights; bf16 rounding equals the bit-level definition on special
values (ties to even, carries into the exponent).
"""
from __future__ import annotations

import numpy as np

NF, NH, NO = 16, 64, 2
SIZES = [("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); float32 out."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init(seed=0, scale=1.0):
    """He-style random weights (float32)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, scale * np.sqrt(2 / NF), (NF, NH)), "b1": np.zeros(NH),
         "W2": rng.normal(0, scale * np.sqrt(2 / NH), (NH, NH)), "b2": np.zeros(NH),
         "W3": rng.normal(0, scale * np.sqrt(1 / NH), (NH, NO)), "b3": np.zeros(NO)}
    return pack(d).astype(np.float32)


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, dtype=np.float64)


def forward(params, X, emulate_bf16=True):
    w = unpack(params)
    e = emulate_bf16
    x = _r(X, e)
    z1 = x @ _r(w["W1"], e) + w["b1"]
    h1 = _r(np.maximum(z1, 0), e)
    z2 = h1 @ _r(w["W2"], e) + w["b2"]
    h2 = _r(np.maximum(z2, 0), e)
    out = h2 @ _r(w["W3"], e) + w["b3"]
    return out, (x, z1, h1, z2, h2)


def loss_and_grads(params, X, Y, M, emulate_bf16=True):
    w = unpack(params)
    e = emulate_bf16
    out, (x, z1, h1, z2, h2) = forward(params, X, e)
    Y = np.asarray(Y, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    n = X.shape[0]
    err = out[:, 0] - Y[:, 0]
    o1 = out[:, 1]
    bce = np.maximum(o1, 0) - o1 * Y[:, 1] + np.log1p(np.exp(-np.abs(o1)))
    loss = float(np.sum(M[:, 0] * err * err + M[:, 1] * bce) / n)
    d3 = np.stack([M[:, 0] * 2 * err / n, M[:, 1] * (1 / (1 + np.exp(-o1)) - Y[:, 1]) / n], axis=1)
    g = {"b3": d3.sum(0)}
    d3r = _r(d3, e)
    g["W3"] = h2.T @ d3r
    dz2 = (d3r @ _r(w["W3"], e).T) * (z2 > 0)
    g["b2"] = dz2.sum(0)
    dz2r = _r(dz2, e)
    g["W2"] = h1.T @ dz2r
    dz1 = (dz2r @ _r(w["W2"], e).T) * (z1 > 0)
    g["b1"] = dz1.sum(0)
    g["W1"] = x.T @ _r(dz1, e)
    return loss, pack(g)


def adam_train(params, X, Y, M, epochs=10, batch=1024, lr=1e-3, seed=0):
    """Reference Adam loop (float64, no rounding) for sanity comparisons."""
    p = np.asarray(params, dtype=np.float64).copy()
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    rng = np.random.default_rng(seed)
    t = 0
    loss = 0.0
    for _ in range(epochs):
        perm = rng.permutation(X.shape[0])
        for b0 in range(0, X.shape[0], batch):
            idx = perm[b0:b0 + batch]
            loss, g = loss_and_grads(p, X[idx], Y[idx], M[idx], emulate_bf16=False)
            t += 1
            m1 = 0.9 * m1 + 0.1 * g
            m2 = 0.999 * m2 + 0.001 * g * g
            p -= lr * (m1 / (1 - 0.9 ** t)) / (np.sqrt(m2 / (1 - 0.999 ** t)) + 1e-8)
    return p.astype(np.float32), loss

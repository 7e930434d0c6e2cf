"""The scheduler network's CPU oracle (oracle/mlp.py), pinned independently of
the GPU kernels: bf16 rounding by its bit-level definition, the forward pass by
hand-set weights, the backward pass by central finite differences."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0, 65504.0,
                  1.9999999], np.float32)
    got = mlp.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3*2^-8 ties between 1+2^-7 and 1+2^-6: even (1+2^-6) wins;
    # 65504 and 1.9999999 round up across a power of two (carry into the exponent)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0, 65536.0, 2.0], np.float32)
    assert np.array_equal(got, want)
    # every result is representable in 8 mantissa bits
    assert np.all((got.view(np.uint32) & 0xFFFF) == 0)


def test_forward_hand_set_weights():
    d = {k: np.zeros(s) for k, s in mlp.SIZES}
    d["W1"][0, 0] = 1.0     # h1[0] = relu(x0)
    d["W1"][1, 1] = -1.0    # h1[1] = relu(-x1)
    d["W2"][0, 0] = 2.0     # h2[0] = 2 h1[0]
    d["W2"][1, 1] = 1.0
    d["b2"][2] = 3.0        # h2[2] = 3
    d["W3"][0, 0] = 1.0
    d["W3"][1, 1] = 1.0
    d["W3"][2, 1] = -0.5
    d["b3"][:] = [0.25, 0.0]
    p = mlp.pack(d)
    X = np.zeros((3, 16))
    X[0, :2] = [1.5, 2.0]
    X[1, :2] = [-1.0, -4.0]
    X[2, :2] = [0.5, 0.0]
    out, _ = mlp.forward(p, X, emulate_bf16=False)
    want = np.array([[3.25, -1.5], [0.25, 4.0 - 1.5], [1.25, -1.5]])
    assert np.allclose(out, want)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_equals_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = mlp.init(seed).astype(np.float64)
    p += rng.normal(0, 0.05, p.shape)  # non-zero biases
    X = rng.normal(0, 1, (37, 16))
    Y = np.stack([rng.uniform(0, 1, 37), (rng.uniform(0, 1, 37) > 0.5)], 1).astype(np.float64)
    M = (rng.uniform(0, 1, (37, 2)) > 0.3).astype(np.float64)
    loss, g = mlp.loss_and_grads(p, X, Y, M, emulate_bf16=False)
    idx = rng.choice(mlp.NPARAMS, 300, replace=False)
    idx = np.concatenate([idx, np.arange(mlp.NPARAMS - 2, mlp.NPARAMS)])  # b3 too
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp.loss_and_grads(q, X, Y, M, emulate_bf16=False)
        q[i] -= 2 * h
        lm, _ = mlp.loss_and_grads(q, X, Y, M, emulate_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_training_reduces_loss_on_a_learnable_task():
    rng = np.random.default_rng(3)
    X = rng.uniform(0, 1, (2048, 16)).astype(np.float32)
    Y = np.stack([0.5 * X[:, 0] + 0.25 * X[:, 3], (X[:, 1] > X[:, 2]).astype(np.float32)], 1)
    M = np.ones_like(Y)
    p0 = mlp.init(0)
    l0, _ = mlp.loss_and_grads(p0, X, Y, M, emulate_bf16=False)
    p1, _ = mlp.adam_train(p0, X, Y, M, epochs=20, batch=256, lr=3e-3)
    l1, _ = mlp.loss_and_grads(p1, X, Y, M, emulate_bf16=False)
    assert l1 < 0.3 * l0

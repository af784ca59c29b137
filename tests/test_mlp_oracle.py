"""Pins of the numpy MLP oracle (oracle/mlp_oracle.py) on the CPU: its
gradient against central finite differences of its own loss, its BF16
rounding against hand-worked values, and its inference form against the
plain forward pass (equal up to BF16 rounding, relu(z)|w| = relu(z|w|))."""
import numpy as np

import mlp_oracle as mo


def rand_problem(seed, n=64):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, mo.N_PARAMS)
    X = rng.normal(0, 1, (n, mo.F))
    X[:, 15] = 1.0
    y = rng.normal(0, 1, n)
    return p, X, y


def test_gradient_matches_finite_differences():
    p, X, y = rand_problem(1)
    loss, g = mo.loss_and_grad(p, X, y)
    h = 1e-6
    rng = np.random.default_rng(2)
    for i in list(rng.choice(mo.N_PARAMS, 60, replace=False)) + [mo.N_PARAMS - 1]:
        e = np.zeros(mo.N_PARAMS)
        e[i] = h
        num = (mo.loss_and_grad(p + e, X, y)[0] - mo.loss_and_grad(p - e, X, y)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * (1 + abs(num)), (i, num, g[i])
    assert abs(loss - 0.5 * np.mean((mo.forward(p, X) - y) ** 2)) < 1e-12


def test_gradient_detects_a_dropped_relu_mask():
    """A plausible slip (no ReLU mask in dL/dW1) changes the gradient."""
    p, X, y = rand_problem(3)
    _, g = mo.loss_and_grad(p, X, y)
    W1, w2, b2 = mo.unpack(p)
    r = b2 + np.maximum(X @ W1, 0) @ w2 - y
    wrong = (X.T @ np.outer(r, w2) / len(y)).ravel()
    assert np.max(np.abs(wrong - g[: mo.F * mo.H])) > 1e-3


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 1 + 2 ** -7 + 2 ** -9], np.float32)
    got = mo.bf16(v)
    assert list(got) == [1.0, 1.0, 1 + 2 ** -6, -2.5, 1 + 2 ** -7]
    x = np.float32(0.1)
    assert abs(mo.bf16(x) - x) <= 2 ** -9 * x


def test_inference_form_equals_forward_up_to_bf16():
    p, X, y = rand_problem(4, n=256)
    W1, w2, _ = mo.unpack(p)
    f = mo.forward(p, X)
    b = mo.infer_bf16(p, X)
    scale = np.maximum(X @ W1, 0) @ np.abs(w2) + np.abs(X) @ np.abs(W1) @ np.abs(w2)
    assert np.all(np.abs(f - b) <= 0.01 * scale + 1e-6)
    # exact when every value is representable in BF16
    p2 = mo.bf16(np.round(p * 8) / 8).astype(np.float64)
    p2[mo.F * mo.H: mo.F * mo.H + mo.H] = np.sign(p2[mo.F * mo.H: -1]) * 0.5
    X2 = np.round(X * 4) / 4
    assert np.allclose(mo.forward(p2, X2), mo.infer_bf16(p2, X2), rtol=0, atol=1e-9)

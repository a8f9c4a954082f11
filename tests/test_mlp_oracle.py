"""The warm-start network's FP64 oracle, pinned without retyping it: GELU by
closed-form values, limits and its symmetry gelu(x) - gelu(-x) = x; the
backward pass by central finite differences of the loss; the layout by a
one-layer network that reduces to least squares."""
import math

import numpy as np

from oracle import mlp_oracle as mo


def test_gelu_closed_forms():
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    x = np.linspace(-6, 6, 121)
    np.testing.assert_allclose(mo.gelu(x) - mo.gelu(-x), x, atol=1e-14)
    # Phi(1) = 0.841344746068543
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.841344746068543) < 1e-14
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) < 1e-12
    h = 1e-6
    num = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(x), num, atol=1e-8)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    ws = mo.widths(5, 7, 3, 4)
    n = sum(a * b + b for a, b in zip(ws[:-1], ws[1:]))
    p = rng.normal(0, 0.5, n)
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 4))
    w = rng.uniform(0, 1, size=(9, 4))
    _, g = mo.loss_grad(p, ws, x, t, w)
    h = 1e-6
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        num = (mo.loss_grad(p + e, ws, x, t, w)[0] - mo.loss_grad(p - e, ws, x, t, w)[0]) / (2 * h)
        assert abs(num - g[i]) < 1e-7 * max(1.0, abs(g[i])), i


def test_linear_network_is_least_squares():
    """With no hidden layer the loss gradient vanishes at the least-squares fit."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(50, 3))
    t = x @ np.array([[1.0, -2.0], [0.5, 0.0], [3.0, 1.0]]) + np.array([0.2, -0.1]) + 0.01 * rng.normal(size=(50, 2))
    X1 = np.hstack([x, np.ones((50, 1))])
    sol = np.linalg.lstsq(X1, t, rcond=None)[0]
    p = np.concatenate([sol[:3].reshape(-1), sol[3]])
    loss, g = mo.loss_grad(p, mo.widths(3, 1, 0, 2), x, t)
    assert np.max(np.abs(g)) < 1e-12
    assert math.isfinite(loss)

"""The float64 NumPy reference of the source-term MLP (oracle/mlp_oracle.py,
the checker of the device's FP32 training kernels) pinned by central
finite differences of its own loss, and its parameter layout by a
hand-built one-layer case."""
import numpy as np

from oracle import mlp_oracle as mo


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    shapes = mo.layer_shapes(6, [8, 8, 8], 3)
    n = sum(fi * fo + fo for fi, fo in shapes)
    p = rng.normal(scale=0.5, size=n)
    x = rng.normal(size=(5, 6))
    t = rng.normal(size=(5, 3))
    loss, g = mo.loss_and_grad(p, shapes, x, t)
    h = 1e-6
    for k in rng.choice(n, size=60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_and_grad(pp, shapes, x, t)[0] - mo.loss_and_grad(pm, shapes, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), k


def test_layout_single_linear_layer():
    """No hidden layer: f(x) = x W + b with W row-major (fan_in x fan_out)
    then b in the parameter vector; the loss gradient is the closed form."""
    shapes = mo.layer_shapes(2, [], 3)
    w = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    b = np.array([0.5, -0.5, 1.0])
    p = np.concatenate([w.ravel(), b])
    x = np.array([[1.0, -1.0], [2.0, 0.5]])
    y, _ = mo.forward(p, shapes, x)
    assert np.allclose(y, x @ w + b)
    t = np.zeros((2, 3))
    loss, g = mo.loss_and_grad(p, shapes, x, t)
    r = x @ w + b
    assert np.isclose(loss, np.mean(r * r))
    assert np.allclose(g[:6].reshape(2, 3), x.T @ (2 * r / r.size))
    assert np.allclose(g[6:], (2 * r / r.size).sum(axis=0))

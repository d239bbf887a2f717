"""Test infrastructure only: float64 NumPy forward and backward of the
source-term MLP (inputs -> 3 hidden tanh layers -> outputs, mean-squared
error), the checker for the device's FP32 training kernels.

The parameter vector is laid out as the device keeps it: for each layer l,
W_l (fan_in x fan_out, row-major) then b_l (fan_out).
"""
from __future__ import annotations

import numpy as np


def layer_shapes(n_in: int, hidden, n_out: int):
    dims = [n_in] + list(hidden) + [n_out]
    return [(dims[k], dims[k + 1]) for k in range(len(dims) - 1)]


def unpack(params: np.ndarray, shapes):
    out, off = [], 0
    for fi, fo in shapes:
        w = params[off:off + fi * fo].reshape(fi, fo)
        off += fi * fo
        b = params[off:off + fo]
        off += fo
        out.append((w, b))
    assert off == len(params)
    return out


def forward(params, shapes, x):
    """Returns the outputs and the hidden activations."""
    layers = unpack(np.asarray(params, dtype=np.float64), shapes)
    h = np.asarray(x, dtype=np.float64)
    acts = [h]
    for k, (w, b) in enumerate(layers):
        z = h @ w + b
        h = np.tanh(z) if k < len(layers) - 1 else z
        acts.append(h)
    return h, acts


def loss_and_grad(params, shapes, x, y):
    """Mean over the batch and the outputs of (f(x) - y)^2 and its gradient
    with respect to the parameter vector."""
    layers = unpack(np.asarray(params, dtype=np.float64), shapes)
    out, acts = forward(params, shapes, x)
    n = out.size
    r = out - np.asarray(y, dtype=np.float64)
    loss = float(np.sum(r * r) / n)
    g = 2.0 * r / n
    grads = []
    for k in range(len(layers) - 1, -1, -1):
        w, _ = layers[k]
        h_in = acts[k]
        grads.append((h_in.T @ g, g.sum(axis=0)))
        if k > 0:
            g = (g @ w.T) * (1.0 - acts[k] * acts[k])
    flat = []
    for gw, gb in reversed(grads):
        flat += [gw.ravel(), gb]
    return loss, np.concatenate(flat)

"""The device's FP32 MLP kernels (rc_mlp_*) against the float64 reference
(oracle/mlp_oracle.py): forward and parameter gradient of the mean-squared
error on seeded inputs (FP32 bar: 1e-4 relative to the largest entry), and
training: the loss falls, and two runs with one seed are bitwise equal."""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from rocketcombustor_amd.surrogate import MLP

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    sys.__stderr__.write("[gpu-test] " + msg + "\n")
    sys.__stderr__.flush()


@pytest.mark.parametrize("n,width,gemm", [(1, 128, "0"), (37, 128, "0"), (300, 128, "0"), (4099, 128, "0"),
                                          (300, 128, "1"), (4099, 128, "1"), (37, 40, "0"), (300, 40, "0")])
def test_forward_and_gradient_match_reference(n, width, gemm, monkeypatch):
    """Width 128 runs the fused MFMA forward kernel and the fused MFMA
    training pass for the gradient (RC_MLP_GEMM=1: the training GEMMs),
    width 40 the GEMM chain for both."""
    monkeypatch.setenv("RC_MLP_GEMM", gemm)
    announce("mlp forward/gradient n=%d width=%d gemm=%s" % (n, width, gemm))
    m = MLP(seed=7, width=width)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 6)).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    p = m.params.astype(np.float64)
    y_ref, _ = mo.forward(p, m.shapes, x.astype(np.float64))
    y = m.forward(x)
    assert np.max(np.abs(y - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = m.gradient(x, t)
    loss_ref, g_ref = mo.loss_and_grad(p, m.shapes, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * np.max(np.abs(g_ref))
    m.close()


@pytest.mark.parametrize("gemm", ["0", "1"], ids=["fused", "gemm-chain"])
def test_training_reduces_loss_and_is_deterministic(gemm, monkeypatch):
    monkeypatch.setenv("RC_MLP_GEMM", gemm)
    announce("mlp training gemm=%s" % gemm)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    t = np.column_stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] * x[:, 3], np.tanh(x[:, 4] - x[:, 5])]).astype(np.float32)
    runs = []
    for _ in range(2):
        m = MLP(seed=3)
        m.set_data(x, t)
        losses = m.train(600, batch=1024, lr=2e-3, seed=9, every=100)
        runs.append((losses, m.params))
        m.close()
    assert runs[0][0][-1] < 0.1 * runs[0][0][0]
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("gemm", ["0", "1"], ids=["fused", "gemm-chain"])
def test_graph_training_equals_direct_launches(gemm, monkeypatch):
    """Training replays captured graphs of 50 steps (the Adam step counter
    and the minibatch stream read on the device); with graphs off the same
    steps launched directly give bitwise the same parameters and losses,
    also across calls that continue the counter."""
    monkeypatch.setenv("RC_MLP_GEMM", gemm)
    announce("mlp training: graph replay vs direct launches gemm=%s" % gemm)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(5000, 6)).astype(np.float32)
    t = np.column_stack([x[:, 0] * x[:, 1], np.sin(x[:, 2]), x[:, 3] - x[:, 4]]).astype(np.float32)
    runs = []
    for no_graph in ("0", "1"):
        monkeypatch.setenv("RC_NO_GRAPH", no_graph)
        m = MLP(seed=3)
        m.set_data(x, t)
        l1 = m.train(120, batch=256, lr=1e-3, seed=9, every=40)
        l2 = m.train(37, batch=256, lr=5e-4, seed=10, every=10)
        runs.append((l1, l2, m.params))
        m.close()
    assert len(runs[0][0]) == 3 and len(runs[0][1]) == 4
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)

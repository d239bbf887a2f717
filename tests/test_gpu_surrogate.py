"""The learned source term in the steady step (rc_set_surrogate): with a
gate that accepts nothing, every cell takes the exact FP64 fallback and the
step equals the exact step bitwise; with a trained network, accepted cells
carry the network's rates, the graph-replayed step equals the eager one,
the counters add up, and restoring the exact source restores the exact
step."""
import sys

import numpy as np
import pytest

from rocketcombustor_amd import native, nozzle, surrogate
from rocketcombustor_amd.solver import SteadyCase

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    sys.__stderr__.write("[gpu-test] " + msg + "\n")
    sys.__stderr__.flush()


@pytest.fixture(scope="module")
def trained():
    announce("surrogate: sample + train")
    case = nozzle.nozzle_case(40, 10)
    raw, omega = surrogate.sample_states([case], steps=[1, 20, 40, 80], per_snapshot=400, seed=1)
    tf = surrogate.Transform.fit(raw, omega)
    m = surrogate.MLP(seed=2)
    m.set_data(tf.x(raw), tf.y(omega))
    losses = m.train(400, batch=512, lr=2e-3, seed=3, every=100)
    assert losses[-1] < losses[0]
    yield m, tf
    m.close()


def exact_run(case, n):
    with SteadyCase(case) as g:
        g.run_steady(n, history=False)
        return g.primitive()


def test_gate_rejecting_all_is_the_exact_step(trained):
    m, tf = trained
    case = nozzle.nozzle_case(40, 10)
    perm, _ = native.order_multicolor(case.mesh.cells)
    case = case.permuted(perm)
    announce("surrogate with an all-rejecting gate")
    ref = exact_run(case, 6)
    with SteadyCase(case) as g:
        g.set_surrogate(m, tf, tol=-1.0)       # |sum| <= negative: nothing accepted
        g.run_steady(6, history=False)
        acc, ev = g.surrogate_stats()
        assert acc == 0 and ev == 6 * case.mesh.ne
        assert np.array_equal(g.primitive(), ref)


def test_trained_network_in_the_step(trained, monkeypatch):
    m, tf = trained
    case = nozzle.nozzle_case(40, 10)
    announce("surrogate accepted in the step, graph vs eager")
    outs = []
    for no_graph in ("0", "1"):
        monkeypatch.setenv("RC_NO_GRAPH", no_graph)
        with SteadyCase(case) as g:
            g.set_surrogate(m, tf, tol=1e9)    # every in-range cell accepted
            g.run_steady(5, history=False)
            acc, ev = g.surrogate_stats()
            assert ev == 5 * case.mesh.ne and 0 < acc <= ev
            outs.append(g.primitive())
            assert np.all(np.isfinite(outs[-1]))
            g.set_surrogate(None)
            g.run_steady(1, history=False)
    assert np.array_equal(outs[0], outs[1])
    # accepted cells carry the network's rates (this briefly trained network's
    # accuracy is not the point here; the measured one is reported by the
    # learned-chemistry bench)
    assert not np.array_equal(outs[0], exact_run(case, 5))

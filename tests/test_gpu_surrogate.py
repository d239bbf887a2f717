"""The learned source term in the steady step (rc_set_surrogate): with a
gate that accepts nothing, every cell takes the exact FP64 fallback and the
step equals the exact step bitwise; with a trained network, accepted cells
carry the network's rates, the graph-replayed step equals the eager one,
the counters add up, and restoring the exact source restores the exact
step; behind the verified gate every cell's source lies within rtol of the
exact FP64 rate; behind the trust gate (a calibrated region, no exact rate)
the device accepts the cells its host restatement accepts.  Time and error
against exact FP64 kinetics at the learned-chemistry workload's size are in
test_gpu_workload_sizes.py."""
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
    losses = surrogate.train_schedule(m, 2000, batch=512, lr=2e-3, seed=3)
    assert losses[-1] < losses[0]
    yield m, tf
    m.close()


def exact_run(case, n):
    with SteadyCase(case) as g:
        g.run_steady(n, history=False)
        return g.primitive()


@pytest.mark.parametrize("nx,ny", [(40, 10), (1, 1)])
def test_gate_rejecting_all_is_the_exact_step(trained, nx, ny, xcd_layout):
    m, tf = trained
    case = nozzle.nozzle_case(nx, ny)
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


def cell_areas(mesh):
    p = mesh.coor[mesh.cells]
    return 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))


def residual_after(g):
    g.timestep()
    g.reconstruction()
    g.flux_sourceterm()
    return g.increment()


@pytest.mark.parametrize("rtol", [0.05, 0.2])
def test_verified_gate_bounds_the_source(trained, rtol):
    m, tf = trained
    case = nozzle.nozzle_case(40, 10)
    announce("surrogate behind the verified gate, rtol %g" % rtol)
    with SteadyCase(case) as g:
        g.run_steady(30, history=False)
        w_exact = g.derived()["production_rate"]
        ref = residual_after(g)
        g.set_surrogate(m, tf, tol=1e9, rtol=rtol)
        got = residual_after(g)
        acc, ev = g.surrogate_stats()
    assert ev == case.mesh.ne and acc > 0
    # only the species rows carry the source; each differs from the exact
    # residual by vol (w - w_exact) with |w - w_exact| <= rtol max(|w_exact|, 1e-3 s_k)
    nsp1 = w_exact.shape[1]
    flow = got.shape[1] - nsp1
    assert np.array_equal(got[:, :flow], ref[:, :flow])
    bound = cell_areas(case.mesh)[:, None] * rtol * np.maximum(np.abs(w_exact), 1e-3 * tf.out_scale)
    d = np.abs(got[:, flow:] - ref[:, flow:])
    assert np.all(d <= bound * (1 + 1e-9) + 1e-12 * np.abs(ref[:, flow:]))
    assert np.count_nonzero(d) > 0          # some cells carry the network's rates


def test_trust_gate_matches_its_host_restatement(trained):
    """The gate with a calibrated trust region (no exact rate formed):
    the device accepts exactly the cells the host restatement accepts, and
    every rejected cell carries the exact FP64 source."""
    m, tf = trained
    case = nozzle.nozzle_case(40, 10)
    announce("surrogate behind the trust gate")
    raw_c, w_c = surrogate.sample_states([case], steps=[10, 30, 60], per_snapshot=400, seed=7)
    trust = surrogate.TrustRegion.calibrate(tf, raw_c, tf.omega(m.forward(tf.x(raw_c))), w_c, dims=(16, 4, 16),
                                            tau=0.05, nmin=2)
    assert 0 < trust.accept.sum() < trust.accept.size
    with SteadyCase(case) as g:
        g.run_steady(30, history=False)
        pa = g.primitive()
        ref = residual_after(g)
        g.set_surrogate(m, tf, tol=1e9, trust=trust)
        got = residual_after(g)
        acc, ev = g.surrogate_stats()
    raw = surrogate.raw_inputs(pa)
    w_net = tf.omega(m.forward(tf.x(raw)))
    ok = surrogate.gate(tf, raw, w_net, None, 1e9, 0.0, trust)
    assert ev == case.mesh.ne
    assert acc == int(ok.sum()) and 0 < acc < ev
    flow = got.shape[1] - w_net.shape[1]
    assert np.array_equal(got[~ok], ref[~ok])
    assert np.array_equal(got[:, :flow], ref[:, :flow])

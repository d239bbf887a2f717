"""Batches of resident cases (rc_batch_*): every member's results are bitwise
the same alone and in a batch, whatever the other members are, including
when members join and leave between calls, and a member's error leaves the
others intact; the mixed batch's members also agree with the C oracle
(<=1e-10 after 60 steps)."""
import sys

import numpy as np
import pytest

from helpers import relerr
from rocketcombustor_amd import native, nozzle
from rocketcombustor_amd.solver import SteadyBatch, SteadyCase

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    sys.__stderr__.write("[gpu-test] " + msg + "\n")
    sys.__stderr__.flush()


def mixed_cases():
    """Different sizes, orders, inlet states, throat ratios, CFL, outlet
    branch and combustion switch (the limiter switch is mixed in the error
    test: without the limiter these nozzles stop within a few steps)."""
    out = [nozzle.nozzle_case(24, 6),
           nozzle.nozzle_case(36, 9, p0=8e6, t0=3400.0, phi=1.2),
           nozzle.nozzle_case(40, 10, combustion=False, cfl=3.0),
           nozzle.nozzle_case(30, 8, p_out_ratio=0.9, subsonic_exit=True)]
    out += [s.build() for s in nozzle.sample_cases(2, cells_lo=800, cells_hi=3000)]
    for k in (1, 4):
        perm, _ = native.order_multicolor(out[k].mesh.cells)
        out[k] = out[k].permuted(perm)
    return out


def alone(case, nsteps, every):
    with SteadyCase(case) as g:
        h = g.run_steady(nsteps, every=every)
        return h, g.primitive(), g.increment()


def test_batch_equals_alone():
    cases = mixed_cases()
    n, every = 60, 10
    announce("batch of %d mixed cases, %d steps" % (len(cases), n))
    ref = [alone(c, n, every) for c in cases]
    gs = [SteadyCase(c) for c in cases]
    with SteadyBatch() as b:
        for g in gs:
            b.add(g)
        info = b.info()
        assert info["members"] == len(cases) and info["cells"] == sum(c.mesh.ne for c in cases)
        hist = b.run_steady(n, every=every)
        for k, g in enumerate(gs):
            assert np.array_equal(hist[:, k], ref[k][0]), k
            assert np.array_equal(g.primitive(), ref[k][1]), k
            assert np.array_equal(g.increment(), ref[k][2]), k
    for g in gs:
        g.close()
    # and against the oracle (every member: these are small)
    for k, c in enumerate(cases):
        o = oracle.OracleCase(c)
        for _ in range(n):
            o.stage()
        assert relerr(ref[k][1], o.get("PA"), -1) < 1e-10, k


def test_batch_dataflow_sweep_equals_alone(monkeypatch):
    """The batch's whole LU-SGS sweep as one dataflow launch (RC_FLOW=1, all
    members, forward and backward) gives every member bitwise what it
    computes alone with level launches."""
    cases = mixed_cases()
    n, every = 30, 10
    announce("batch dataflow sweep, %d mixed cases, %d steps" % (len(cases), n))
    monkeypatch.setenv("RC_FLOW", "0")
    ref = [alone(c, n, every) for c in cases]
    gs = [SteadyCase(c) for c in cases]
    monkeypatch.setenv("RC_FLOW", "1")
    try:
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            hist = b.run_steady(n, every=every)
            for k, g in enumerate(gs):
                assert np.array_equal(hist[:, k], ref[k][0]), k
                assert np.array_equal(g.primitive(), ref[k][1]), k
    finally:
        monkeypatch.delenv("RC_FLOW")
        for g in gs:
            g.close()


def test_batch_next_level_residual_equals_alone(monkeypatch):
    """Batch forward levels that form the next level's residual in extra
    thread blocks (RC_NEXTRES=1, every level; members with different level
    counts) give every member bitwise what it computes alone with each level
    forming its own residual (RC_NEXTRES=0)."""
    cases = mixed_cases()
    n, every = 30, 10
    announce("batch next-level residual, %d mixed cases, %d steps" % (len(cases), n))
    monkeypatch.setenv("RC_NEXTRES", "0")
    ref = [alone(c, n, every) for c in cases]
    gs = [SteadyCase(c) for c in cases]
    monkeypatch.setenv("RC_NEXTRES", "1")
    try:
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            hist = b.run_steady(n, every=every)
            for k, g in enumerate(gs):
                assert np.array_equal(hist[:, k], ref[k][0]), k
                assert np.array_equal(g.primitive(), ref[k][1]), k
    finally:
        monkeypatch.delenv("RC_NEXTRES")
        for g in gs:
            g.close()


def test_continuous_batching_join_and_leave():
    """A, B from the start; C joins after 30 steps; A leaves after 50; each
    ends bitwise where it ends alone after the same number of steps."""
    ca, cb, cc = mixed_cases()[:3]
    announce("continuous batching join/leave")
    ga, gb, gc = SteadyCase(ca), SteadyCase(cb), SteadyCase(cc)
    with SteadyBatch() as b:
        b.add(ga)
        b.add(gb)
        b.run_steady(30, history=False)
        b.add(gc)
        b.run_steady(20, history=False)
        b.remove(ga)
        b.run_steady(25, history=False)
        assert b.info()["members"] == 2
    for g, c, steps in ((ga, ca, 50), (gb, cb, 75), (gc, cc, 45)):
        assert np.array_equal(g.primitive(), alone(c, steps, 0)[1])
        g.close()


@pytest.mark.parametrize("every", [1, 10])
def test_member_error_is_isolated(every):
    """The no-limiter case stops at step 5 (as the reference does); the
    error names the member, and the healthy member continues to the same
    state as alone (reports every step, or every 10 steps with the steps
    between them replayed as graphs of several steps)."""
    bad = nozzle.nozzle_case(24, 6, limit_id=0)
    good = nozzle.nozzle_case(30, 8)
    announce("batch member error")
    gbad, ggood = SteadyCase(bad), SteadyCase(good)
    with SteadyBatch() as b:
        b.add(ggood)
        b.add(gbad)
        with pytest.raises(native.RcError) as ei:
            b.run_steady(20, every=every)
        assert "case 1" in str(ei.value)
        done = ei.value.done
        assert ei.value.member == 1 and 1 <= done < 20
        b.remove(gbad)
        b.run_steady(20 - done, history=False)
    assert np.array_equal(ggood.primitive(), alone(good, 20, 0)[1])
    gbad.close()
    ggood.close()


def test_batch_graph_replay_matches_eager(monkeypatch):
    """The batch step as a replayed graph and as eager launches agree bitwise
    (several replays with host work between them)."""
    cases = mixed_cases()[:3]
    announce("batch graph vs eager")
    outs = []
    for no_graph in ("0", "1"):
        monkeypatch.setenv("RC_NO_GRAPH", no_graph)
        gs = [SteadyCase(c) for c in cases]
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            for _ in range(4):
                b.run_steady(7, history=False)
                _ = [g.primitive() for g in gs]
            outs.append([g.primitive() for g in gs])
        for g in gs:
            g.close()
    for a, e in zip(*outs):
        assert np.array_equal(a, e)


def test_batch_multi_step_graph_matches_single_steps(monkeypatch):
    """The batch's runs of steps between residual reports as graphs of
    several steps equal one graph per step and eager launches bitwise,
    members' states and every reported residual row."""
    cases = mixed_cases()[:3]
    announce("batch multi-step graph vs one graph per step vs eager")
    outs = []
    for graph_steps, no_graph in ((None, "0"), ("1", "0"), (None, "1")):
        monkeypatch.setenv("RC_NO_GRAPH", no_graph)
        if graph_steps is None:
            monkeypatch.delenv("RC_GRAPH_STEPS", raising=False)
        else:
            monkeypatch.setenv("RC_GRAPH_STEPS", graph_steps)
        gs = [SteadyCase(c) for c in cases]
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            h1 = b.run_steady(25, every=9)
            _ = [g.primitive() for g in gs]
            h2 = b.run_steady(17, history=False)
            outs.append((h1, [g.primitive() for g in gs], [g.increment() for g in gs]))
        for g in gs:
            g.close()
    for other in outs[1:]:
        assert np.array_equal(outs[0][0], other[0])
        for a, e in zip(outs[0][1], other[1]):
            assert np.array_equal(a, e)
        for a, e in zip(outs[0][2], other[2]):
            assert np.array_equal(a, e)


def test_batch_with_periodic_members_equals_alone():
    """Members with periodic boundary pairs (their TimeStep sums over the
    periodic faces too) beside ordinary ones: each member bitwise its alone
    result, residual rows included."""
    px = nozzle.periodic_channel_case(30, 8, direction="x")
    py = nozzle.periodic_channel_case(24, 6, direction="y", p_out_ratio=0.5)
    perm, _ = native.order_multicolor(py.mesh.cells)
    cases = [px, nozzle.nozzle_case(24, 6), py.permuted(perm)]
    n, every = 20, 5
    announce("batch with periodic members, %d cases" % len(cases))
    ref = [alone(c, n, every) for c in cases]
    gs = [SteadyCase(c) for c in cases]
    with SteadyBatch() as b:
        for g in gs:
            b.add(g)
        h = b.run_steady(n, every=every)
        for k, g in enumerate(gs):
            assert np.array_equal(h[:, k], ref[k][0]), k
            assert np.array_equal(g.primitive(), ref[k][1]), k
            assert np.array_equal(g.increment(), ref[k][2]), k
    for g in gs:
        g.close()


def test_batch_fused_top_level_equals_separate_backward(monkeypatch):
    """The batch's highest level running its backward sweep in its forward
    launch equals a backward launch of its own bitwise (members with fewer
    levels than the batch included)."""
    cases = mixed_cases()
    announce("batch fused top level vs separate backward, %d mixed cases" % len(cases))
    outs = []
    for fuse in ("0", None):
        if fuse is None:
            monkeypatch.delenv("RC_TOPFUSE", raising=False)
        else:
            monkeypatch.setenv("RC_TOPFUSE", fuse)
        gs = [SteadyCase(c) for c in cases]
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            h = b.run_steady(14, every=5)
            outs.append((h, [g.primitive() for g in gs], [g.increment() for g in gs]))
        for g in gs:
            g.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    for a, e in zip(outs[0][1] + outs[0][2], outs[1][1] + outs[1][2]):
        assert np.array_equal(a, e)


@pytest.mark.parametrize("shape", [(64, 64, 0), (128, 256, 1), (64, 128, 1), (256, 64, -1)])
def test_launch_shape_does_not_change_results(shape):
    """The autotuner's knob (threads per block, XCD banding) re-cuts the work
    lists only: every member stays bitwise its alone result, also when the
    shape changes between calls."""
    cases = mixed_cases()
    n, every = 20, 5
    announce("batch launch shape %s, %d mixed cases" % (shape, len(cases)))
    ref = [alone(c, n, every) for c in cases]
    gs = [SteadyCase(c) for c in cases]
    with SteadyBatch() as b:
        for g in gs:
            b.add(g)
        b.set_launch(*shape)
        h1 = b.run_steady(10, every=every)
        b.set_launch(64, 64, 1)
        h2 = b.run_steady(10, every=every)
        for k, g in enumerate(gs):
            assert np.array_equal(h1[:, k], ref[k][0][:3]), k
            assert np.array_equal(h2[1:, k], ref[k][0][3:]), k
            assert np.array_equal(g.primitive(), ref[k][1]), k
    for g in gs:
        g.close()


def test_launch_shape_rejects_bad_values():
    with SteadyBatch() as b:
        for bad in [(32, 256, 1), (256, 512, 1), (256, 256, 2)]:
            with pytest.raises(native.RcError):
                b.set_launch(*bad)

"""The persistent step (RC_PERSIST=n: n steady steps per cooperative launch,
grid barriers between the phases) against the graph-replayed step: the
state, the increments and every reported residual row are equal bit for bit,
for one case and for a batch, with few and many thread blocks (RC_PERSIST_GRID
caps the grid, so each block loops over many items of every phase), and for
runs whose length is not a multiple of the steps per launch."""
import sys

import numpy as np
import pytest

from rocketcombustor_amd import native, nozzle
from rocketcombustor_amd.solver import SteadyBatch, SteadyCase

pytestmark = pytest.mark.gpu

# levels the persistent step supports (more: the graph path runs instead)
MAX_LEVELS = 32


def announce(msg: str) -> None:
    sys.__stderr__.write("[gpu-test] " + msg + "\n")
    sys.__stderr__.flush()


def multicolor(case):
    perm, _ = native.order_multicolor(case.mesh.cells)
    return case.permuted(perm)


def run_case(case, monkeypatch, persist, grid=None):
    if persist:
        monkeypatch.setenv("RC_PERSIST", str(persist))
    else:
        monkeypatch.delenv("RC_PERSIST", raising=False)
    if grid:
        monkeypatch.setenv("RC_PERSIST_GRID", str(grid))
    else:
        monkeypatch.delenv("RC_PERSIST_GRID", raising=False)
    with SteadyCase(case) as g:
        assert g.nlev <= MAX_LEVELS
        hist = [g.run_steady(1), g.run_steady(30, every=10)]
        first = (g.primitive(), g.increment())
        hist.append(g.run_steady(13, every=5))
        return hist, first, (g.primitive(), g.increment(), g.conserved())


@pytest.mark.parametrize("nx,ny,persist,grid", [(24, 6, 8, None), (150, 40, 8, None), (150, 40, 3, 2),
                                                (150, 40, 1, 7), (1000, 200, 8, None), (1000, 200, 8, 64)])
def test_persistent_step_equals_graph(nx, ny, persist, grid, monkeypatch):
    case = multicolor(nozzle.nozzle_case(nx, ny))
    announce("persistent step %dx%d, %d steps per launch, grid cap %s" % (nx, ny, persist, grid))
    ref = run_case(case, monkeypatch, 0)
    got = run_case(case, monkeypatch, persist, grid)
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    for sa, sb in zip(ref[1:], got[1:]):
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("direction", ["x", "y"])
def test_persistent_step_equals_graph_periodic(direction, monkeypatch):
    """Periodic boundary pairs through the persistent step."""
    case = multicolor(nozzle.periodic_channel_case(60, 16, direction=direction, p_out_ratio=0.5))
    announce("persistent step, periodic channel (%s)" % direction)
    ref = run_case(case, monkeypatch, 0)
    got = run_case(case, monkeypatch, 8, 16)
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    for sa, sb in zip(ref[1:], got[1:]):
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("grid", [None, 5])
def test_batch_persistent_step_equals_graph(grid, monkeypatch):
    """Mixed members (sizes, inlet states, combustion switch, outlet branch),
    every one multicolour ordered."""
    cases = [multicolor(c) for c in (nozzle.nozzle_case(24, 6),
                                     nozzle.nozzle_case(36, 9, p0=8e6, t0=3400.0, phi=1.2),
                                     nozzle.nozzle_case(40, 10, combustion=False, cfl=3.0),
                                     nozzle.nozzle_case(30, 8, p_out_ratio=0.9, subsonic_exit=True))]
    announce("batch persistent step, %d members, grid cap %s" % (len(cases), grid))
    outs = []
    for persist in (0, 8):
        if persist:
            monkeypatch.setenv("RC_PERSIST", str(persist))
        else:
            monkeypatch.delenv("RC_PERSIST", raising=False)
        if grid:
            monkeypatch.setenv("RC_PERSIST_GRID", str(grid))
        else:
            monkeypatch.delenv("RC_PERSIST_GRID", raising=False)
        gs = [SteadyCase(c) for c in cases]
        with SteadyBatch() as b:
            for g in gs:
                b.add(g)
            assert b.info()["levels"] <= MAX_LEVELS
            h1 = b.run_steady(25, every=9)
            h2 = b.run_steady(11, history=False)
            assert h2 is None
            outs.append((h1, [g.primitive() for g in gs], [g.increment() for g in gs]))
        for g in gs:
            g.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    for a, e in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(a, e)
    for a, e in zip(outs[0][2], outs[1][2]):
        assert np.array_equal(a, e)

"""Serving on the GPU: requests arriving while others run are admitted and
stepped in the same batch; every served result is bitwise the one the same
case reaches alone in the same number of steps, and the first two agree with
the C oracle run for those steps (<=1e-8 over up to 1200 steps)."""
import sys

import numpy as np
import pytest

from helpers import relerr
from rocketcombustor_amd import native, serving
from rocketcombustor_amd.solver import SteadyCase

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    sys.__stderr__.write("[gpu-test] serving 8 small requests\n")
    sys.__stderr__.flush()
    reqs = serving.poisson_requests(8, rate=40.0, cells_lo=600, cells_hi=3000)
    final = {}
    srv = serving.Server(round_steps=40, orders=1.0, max_steps=1200, workers=2,
                         on_finish=lambda r, case: final.__setitem__(r.rid, case.primitive()))
    rep = srv.run(reqs, time_limit=120.0)
    assert rep["unfinished"] == 0 and rep["failed"] == 0
    assert rep["converged"] + rep["capped"] == len(reqs)
    assert rep["converged"] >= 1
    assert rep["mean_members"] > 1.0          # requests overlapped on the device
    for r in reqs:
        with serving.build_case(r.spec) as g:
            g.run_steady(r.steps, history=False)
            assert np.array_equal(g.primitive(), final[r.rid]), r.rid
    # the first two served results against the oracle run for the same steps
    for r in reqs[:2]:
        case = r.spec.build()
        perm, _ = native.order_multicolor(case.mesh.cells)
        o = oracle.OracleCase(case.permuted(perm))
        for _ in range(r.steps):
            o.stage()
        assert relerr(final[r.rid], o.get("PA"), -1) < 1e-8, r.rid
    assert np.isfinite(rep["p50_s"]) and rep["p50_s"] <= rep["p99_s"]

"""Host logic of the serving loop with stand-in cases (no GPU): admission
by the memory budget, retirement on convergence, step cap, member failure,
the latency report, and the seeded request stream."""
import math

import numpy as np
import pytest

from rocketcombustor_amd import serving
from rocketcombustor_amd.native import NVAR, RcError


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeCase:
    def __init__(self, spec):
        self.ne = spec.cells
        self.device_bytes = int(900 * spec.cells)
        self.tau = 50.0 * math.log10(spec.cells)     # steps per e-fold of the residual
        self.steps = 0
        self.closed = False

    def close(self):
        self.closed = True

    def run_steady(self, n, every=0, history=True):
        """The member's first step alone (taken before it joins the batch)."""
        h = np.zeros((1, NVAR + 1))
        h[0, 4:NVAR] = math.exp(-self.steps / self.tau)
        self.steps += n
        return h


class FakeBatch:
    """Stand-in for SteadyBatch: RoRes decays as exp(-steps/tau); the clock
    advances with the cell-steps done; optionally one case fails."""

    def __init__(self, clock, budget, fail_cells=None):
        self.members, self.clock, self.budget = [], clock, budget
        self.peak_bytes = 0
        self.fail_cells = fail_cells

    def add(self, case):
        self.members.append(case)
        self.peak_bytes = max(self.peak_bytes, sum(c.device_bytes for c in self.members))
        assert self.peak_bytes <= self.budget

    def remove(self, case):
        self.members.remove(case)

    def run_steady(self, n, every=0, history=True):
        for k, c in enumerate(self.members):
            if self.fail_cells is not None and c.ne == self.fail_cells and c.steps >= n:
                e = RcError(2, "case %d: temperature" % k)
                e.done, e.member = 3, k
                for m in self.members:
                    m.steps += 3
                raise e
        h = np.zeros((2, len(self.members), NVAR + 1))
        for k, c in enumerate(self.members):
            h[0, k, 4:NVAR] = math.exp(-c.steps / c.tau)
            c.steps += n
            h[1, k, 4:NVAR] = math.exp(-c.steps / c.tau)
        self.clock.t += 1e-9 * n * sum(c.ne for c in self.members) + 1e-4
        return h


def run(reqs, budget=50e6, **kw):
    clock = FakeClock()
    batch = FakeBatch(clock, budget, kw.pop("fail_cells", None))
    srv = serving.Server(hbm_budget=budget, round_steps=50, batch=batch, make_case=FakeCase, clock=clock,
                         workers=1, **kw)
    return srv.run(reqs), reqs, batch


def test_request_stream_is_seeded_poisson():
    a = serving.poisson_requests(2000, rate=20.0)
    b = serving.poisson_requests(2000, rate=20.0)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 20.0) < 0.005 and np.all(gaps > 0)
    cells = np.array([r.cells for r in a])
    assert cells.min() >= 4000 and cells.max() <= 4.2e5
    # log-uniform: the median is near the geometric mean of the range
    assert 3.0e4 < np.median(cells) < 6.0e4


def test_admission_budget_and_convergence():
    reqs = [serving.Request(rid=k, arrival=0.0, spec=s) for k, s in
            enumerate(serving.nozzle.sample_cases(12, cells_lo=5e3, cells_hi=4e4))]
    for r in reqs:
        r.cells = r.spec.cells
    rep, reqs, batch = run(reqs, budget=60e6)
    assert rep["converged"] == 12 and rep["failed"] == rep["capped"] == 0
    assert batch.peak_bytes <= 60e6
    assert batch.members == []
    for r in reqs:
        # one step alone, then converged at the first round end past
        # ln(100)*tau steps
        tau = 50.0 * math.log10(r.cells)
        assert r.steps == 1 + 50 * math.ceil((math.log(100.0) * tau - 1) / 50)
        assert r.res <= 1e-2 * r.res0
    assert rep["p50_s"] <= rep["p99_s"]


def test_step_cap_failure_and_rejection():
    specs = serving.nozzle.sample_cases(4, cells_lo=5e3, cells_hi=4e4)
    reqs = [serving.Request(rid=k, arrival=0.0, spec=s, cells=s.cells) for k, s in enumerate(specs)]
    big = serving.nozzle.CaseSpec(nx=2000, ny=400, p0=5e6, t0=3000.0, phi=1.0, throat_ratio=0.4)
    reqs.append(serving.Request(rid=9, arrival=0.0, spec=big, cells=big.cells))
    rep, reqs, _ = run(reqs, budget=60e6, max_steps=100, fail_cells=specs[0].cells)
    st = {r.rid: r.status for r in reqs}
    assert st[9] == "rejected"
    assert st[0] == "failed"
    assert all(st[k] == "capped" for k in (1, 2, 3))
    assert rep["converged"] == 0 and rep["capped"] == 3 and rep["failed"] == 1 and rep["rejected"] == 1
    # the members that ran on past the failed round keep their own step-1
    # baseline (RoRes of a fresh stand-in: three species rows of 1.0)
    assert all(r.res0 == 3.0 for r in reqs if r.rid in (1, 2, 3))

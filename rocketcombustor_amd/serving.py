"""Serving: a request loop over resident steady cases with continuous
batching (no reference counterpart; the reference solves one case per run).

Requests (nozzle/chamber cases) arrive over time.  A request is admitted
when the device-memory budget has room for it: its case is built on a host
worker thread (mesh, sweep order, device set-up through the C-ABI) while the
GPU keeps stepping the cases already resident.  The resident cases form one
:class:`SteadyBatch`: every round runs ``round_steps`` steady steps of all of
them with the same kernel launches.  After each round a case whose RoRes
(the species part of step_residual, the quantity the reference monitors,
src/residual.f90:79) has dropped ``orders`` orders below its first-step value
is converged; it leaves the batch and frees its memory, and waiting requests
take its place.  A case that reaches ``max_steps`` leaves as capped, one whose
step fails (the reference would stop) leaves as failed; a request larger than
the whole budget is rejected.

Every member computes exactly what it computes alone (the batch kernels are
per-case gathers), so a served result equals the same case run alone for the
same number of steps.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import math
import time
from collections import deque
from typing import Callable, List, Optional

import numpy as np

from . import nozzle
from .native import NVAR, RcError

# device bytes per cell of a resident case, for admission before the case
# exists (measured set-up sizes are about 900-1000 B per triangle)
BYTES_PER_CELL_ESTIMATE = 1100.0


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                  # seconds after the start of the run
    spec: nozzle.CaseSpec
    admitted: float = math.nan      # build submitted
    started: float = math.nan       # first step
    finished: float = math.nan
    steps: int = 0
    status: str = "pending"         # pending, building, running, converged, capped, failed, rejected
    res0: float = math.nan          # RoRes at the first step
    res: float = math.nan           # RoRes at the last report
    device_bytes: int = 0
    cells: int = 0
    error: str = ""

    @property
    def latency(self) -> float:
        """Time to converged: arrival to leaving the batch."""
        return self.finished - self.arrival


def poisson_requests(n: int, rate: float, seed: int = 4, cells_lo: float = 5e3, cells_hi: float = 4e5,
                     first: int = 0) -> List[Request]:
    """``n`` requests with exponential inter-arrival times of mean 1/rate
    (SplitMix64 stream ``seed``) and case parameters of the seeded family
    (:func:`nozzle.sample_cases` with the same seed)."""
    u = nozzle.uniform01(seed + 1000, np.arange(first, first + n))
    gaps = -np.log1p(-u) / rate
    t = np.cumsum(gaps)
    specs = nozzle.sample_cases(n, seed=seed, cells_lo=cells_lo, cells_hi=cells_hi, first=first)
    return [Request(rid=first + k, arrival=float(t[k]), spec=specs[k], cells=specs[k].cells) for k in range(n)]


def roress(res_row: np.ndarray) -> float:
    """RoRes of one step_residual row: the sum of the species increments'
    norms (src/residual.f90:79)."""
    return float(np.sum(res_row[4:NVAR]))


def build_case(spec: nozzle.CaseSpec, device: int = 0, block: int = 4):
    """Host + device set-up of one request: seeded nozzle case in the
    block-multicolour sweep order, resident on ``device``."""
    from . import native
    from .solver import SteadyCase
    case = spec.build()
    perm, _ = native.order_multicolor(case.mesh.cells, block=block)
    return SteadyCase(case.permuted(perm), device=device)


class Server:
    """Continuous-batching request loop.  ``make_case(spec)`` builds a
    resident case (with ``.ne`` and ``.device_bytes``, ``.close()``) and
    ``batch`` is a :class:`SteadyBatch`-like object (``add``, ``remove``,
    ``run_steady``); both default to the GPU ones."""

    def __init__(self, device: int = 0, hbm_budget: float = 288e9, round_steps: int = 200, orders: float = 2.0,
                 max_steps: int = 20000, workers: int = 4, make_case: Optional[Callable] = None, batch=None,
                 clock: Callable[[], float] = time.perf_counter, max_members: int = 4096,
                 on_finish: Optional[Callable] = None):
        if batch is None:
            from .solver import SteadyBatch
            batch = SteadyBatch(device)
        self.batch = batch
        self.make_case = make_case or (lambda spec: build_case(spec, device))
        self.budget = float(hbm_budget)
        self.round_steps = int(round_steps)
        self.drop = 10.0 ** (-orders)
        self.max_steps = int(max_steps)
        self.workers = int(workers)
        self.clock = clock
        self.max_members = int(max_members)
        self.on_finish = on_finish      # on_finish(request, case) before a case is freed

    def run(self, requests: List[Request], time_limit: Optional[float] = None) -> dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = deque(reqs)
        building = {}                 # future -> request
        active = []                   # (request, case) in batch order
        reserved = 0.0                # bytes of building + resident cases
        rounds = cell_steps = 0
        members_sum = 0
        t0 = self.clock()
        pool = cf.ThreadPoolExecutor(max_workers=self.workers)
        try:
            while pending or building or active:
                now = self.clock() - t0
                if time_limit is not None and now > time_limit:
                    break
                # admission: arrived requests while the budget has room
                while pending and pending[0].arrival <= now and len(building) + len(active) < self.max_members:
                    r = pending[0]
                    est = BYTES_PER_CELL_ESTIMATE * r.cells
                    if est > self.budget:
                        pending.popleft()
                        r.status, r.finished = "rejected", now
                        continue
                    if reserved + est > self.budget:
                        break
                    pending.popleft()
                    reserved += est
                    r.device_bytes = int(est)
                    r.admitted, r.status = now, "building"
                    building[pool.submit(self.make_case, r.spec)] = r
                # built cases join the batch
                for fut in [f for f in building if f.done()]:
                    r = building.pop(fut)
                    try:
                        case = fut.result()
                    except RcError as e:
                        reserved -= r.device_bytes
                        r.status, r.finished, r.error = "failed", self.clock() - t0, str(e)
                        continue
                    try:
                        # the first step alone: the member's baseline is its own
                        # step-1 RoRes, whatever its first batched round does
                        h = case.run_steady(1, every=1)
                    except RcError as e:
                        reserved -= r.device_bytes
                        r.steps = 1
                        r.status, r.finished, r.error = "failed", self.clock() - t0, str(e)
                        case.close()
                        continue
                    r.steps, r.res0 = 1, roress(h[0])
                    r.res = r.res0
                    reserved += case.device_bytes - r.device_bytes
                    r.device_bytes, r.cells = int(case.device_bytes), int(case.ne)
                    self.batch.add(case)
                    active.append((r, case))
                    r.status, r.started = "running", self.clock() - t0
                if not active:
                    # nothing resident: wait for a build or the next arrival
                    if building:
                        cf.wait(list(building), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                    elif pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - (self.clock() - t0))))
                    continue
                # one round of batched steps
                n = self.round_steps
                done_steps, failed, err = n, None, ""
                try:
                    hist = self.batch.run_steady(n, every=n)
                except RcError as e:
                    hist, done_steps, failed, err = None, getattr(e, "done", 0), getattr(e, "member", -1), str(e)
                rounds += 1
                members_sum += len(active)
                cell_steps += done_steps * sum(r.cells for r, _ in active)
                now = self.clock() - t0
                leave = []
                for k, (r, case) in enumerate(active):
                    r.steps += done_steps
                    if k == failed:
                        r.status, r.error = "failed", err
                        leave.append(k)
                        continue
                    if hist is None:
                        continue
                    r.res = roress(hist[-1, k])
                    if r.res <= self.drop * r.res0:
                        r.status = "converged"
                        leave.append(k)
                    elif r.steps >= self.max_steps:
                        r.status = "capped"
                        leave.append(k)
                for k in reversed(leave):
                    r, case = active.pop(k)
                    r.finished = now
                    self.batch.remove(case)
                    if self.on_finish is not None:
                        self.on_finish(r, case)
                    case.close()
                    reserved -= r.device_bytes
        finally:
            for fut, r in building.items():
                try:
                    fut.result().close()
                except Exception:
                    pass
            for r, case in active:
                self.batch.remove(case)
                case.close()
            pool.shutdown(wait=True)
        elapsed = self.clock() - t0
        return self.report(reqs, elapsed, rounds, cell_steps, members_sum)

    def report(self, reqs: List[Request], elapsed: float, rounds: int, cell_steps: int, members_sum: int) -> dict:
        conv = [r for r in reqs if r.status == "converged"]
        lat = np.array([r.latency for r in conv]) if conv else np.array([math.nan])
        offered = len(reqs) / max(reqs[-1].arrival, 1e-9) if reqs else 0.0
        return {
            "requests": len(reqs),
            "converged": len(conv),
            "capped": sum(r.status == "capped" for r in reqs),
            "failed": sum(r.status == "failed" for r in reqs),
            "rejected": sum(r.status == "rejected" for r in reqs),
            "unfinished": sum(r.status in ("pending", "building", "running") for r in reqs),
            "elapsed_s": elapsed,
            "cases_per_hour": 3600.0 * len(conv) / elapsed if elapsed > 0 else 0.0,
            "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)),
            "offered_req_per_s": offered,
            "served_req_per_s": len(conv) / elapsed if elapsed > 0 else 0.0,
            "rounds": rounds,
            "mean_members": members_sum / rounds if rounds else 0.0,
            "cell_updates_per_s": cell_steps / elapsed if elapsed > 0 else 0.0,
            "steps_to_converge_mean": float(np.mean([r.steps for r in conv])) if conv else math.nan,
            "failures": [{"rid": r.rid, "cells": r.cells, "steps": r.steps, "error": r.error}
                         for r in reqs if r.status == "failed"],
        }


def serve_sequential(requests: List[Request], round_steps: int = 200, orders: float = 2.0, max_steps: int = 20000,
                     device: int = 0, time_limit: Optional[float] = None,
                     clock: Callable[[], float] = time.perf_counter) -> dict:
    """The plain path for comparison: requests served one at a time in
    arrival order, each case alone on the device through the single-case
    step loop (rc_run_steady), with the same convergence test."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    drop = 10.0 ** (-orders)
    cell_steps = 0
    t0 = clock()
    for r in reqs:
        now = clock() - t0
        if time_limit is not None and now > time_limit:
            break
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.admitted = clock() - t0
        with build_case(r.spec, device) as g:
            r.cells, r.device_bytes, r.status = g.ne, g.device_bytes, "running"
            r.started = clock() - t0
            try:
                while True:
                    h = g.run_steady(round_steps, every=round_steps)
                    r.steps += round_steps
                    cell_steps += round_steps * g.ne
                    if math.isnan(r.res0):
                        r.res0 = roress(h[0])
                    r.res = roress(h[-1])
                    if r.res <= drop * r.res0:
                        r.status = "converged"
                        break
                    if r.steps >= max_steps:
                        r.status = "capped"
                        break
            except RcError as e:
                r.status, r.error = "failed", str(e)
        r.finished = clock() - t0
    elapsed = clock() - t0
    srv = Server.__new__(Server)
    return Server.report(srv, reqs, elapsed, 0, cell_steps, 0)

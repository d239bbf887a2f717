"""Autotuner for the batched steady step (no reference counterpart: the
reference steps one case with serial loops).

Two knobs are searched by timing graph-replayed steps of the batch itself:

* the launch shape of the batch (``rc_batch_set_launch``): threads per
  block of the cell / face / ghost kernels, threads per block of the LU-SGS
  level kernels (four lanes per run), and XCD banding of the work lists.
  Results do not depend on it: each cell and face is computed by one thread
  in the same order whatever the shape, so a tuned batch is bitwise the
  untuned one;
* the colouring granularity of the cases' block-multicolour order
  (``rc_mesh_order_multicolor``'s block size): it sets how many sweep levels
  a step launches and how much Gauss-Seidel coupling each block keeps.  It
  changes the sweep order, so the transient (not the converged field); block
  1 (per-cell colouring) stalls the steady nozzle and is not offered.

The objective is wall time per graph-replayed step of the whole batch.
"""
from __future__ import annotations

import dataclasses
import json
import time
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

THREADS = (64, 128, 256)
SHAPES: Tuple[Tuple[int, int, int], ...] = tuple((c, r, x) for c in THREADS for r in THREADS for x in (1, 0))
BLOCKS = (2, 4, 8)


@dataclasses.dataclass
class Trial:
    block: int
    cell_threads: int
    run_threads: int
    xcd: int
    ms_per_step: float
    levels: int


def time_steps(batch, steps: int, warmup: int, timer: Callable[[], float] = time.perf_counter) -> float:
    """Milliseconds per step of ``steps`` graph-replayed steps after
    ``warmup`` (which also rebuilds the work lists and captures the graph)."""
    if warmup > 0:
        batch.run_steady(warmup, history=False)
    t0 = timer()
    batch.run_steady(steps, history=False)
    return 1e3 * (timer() - t0) / max(steps, 1)


def tune_launch(batch, shapes: Sequence[Tuple[int, int, int]] = SHAPES, steps: int = 10, warmup: int = 2,
                timer: Callable[[], float] = time.perf_counter, block: int = 0) -> Tuple[Trial, List[Trial]]:
    """Times every launch shape on ``batch`` (a :class:`SteadyBatch`) and
    leaves the fastest set; returns it and all trials.  The timed steps
    advance the members like any other steps."""
    levels = batch.info()["levels"]
    trials = []
    for c, r, x in shapes:
        batch.set_launch(c, r, x)
        trials.append(Trial(block, c, r, x, time_steps(batch, steps, warmup, timer), levels))
    best = min(trials, key=lambda t: t.ms_per_step)
    batch.set_launch(best.cell_threads, best.run_threads, best.xcd)
    return best, trials


def tune(specs: Iterable, blocks: Sequence[int] = BLOCKS, shapes: Sequence[Tuple[int, int, int]] = SHAPES,
         steps: int = 10, warmup: int = 2, device: int = 0, make_case: Optional[Callable] = None,
         make_batch: Optional[Callable] = None, timer: Callable[[], float] = time.perf_counter,
         workers: int = 1) -> dict:
    """Colouring granularity and launch shape for a workload: for each block
    size the sample cases ``specs`` are built in that order (``workers``
    host threads), batched, and every launch shape timed.
    ``make_case(spec, block)`` and ``make_batch()`` default to the device
    ones.  Returns the best setting with every trial."""
    import concurrent.futures as cf
    specs = list(specs)
    if make_case is None:
        from .serving import build_case

        def make_case(spec, block):
            return build_case(spec, device=device, block=block)
    if make_batch is None:
        from .solver import SteadyBatch

        def make_batch():
            return SteadyBatch(device=device)
    trials: List[Trial] = []
    for block in blocks:
        if workers > 1:
            with cf.ThreadPoolExecutor(max_workers=workers) as pool:
                cases = list(pool.map(lambda s: make_case(s, block), specs))
        else:
            cases = [make_case(s, block) for s in specs]
        b = make_batch()
        try:
            for g in cases:
                b.add(g)
            _, tr = tune_launch(b, shapes, steps, warmup, timer, block=block)
            trials += tr
            for g in cases:
                b.remove(g)
        finally:
            b.close()
            for g in cases:
                g.close()
    best = min(trials, key=lambda t: t.ms_per_step)
    default = [t for t in trials if (t.block, t.cell_threads, t.run_threads, t.xcd) == (4, 256, 64, 1)]
    return {"block": best.block, "cell_threads": best.cell_threads, "run_threads": best.run_threads,
            "xcd": best.xcd, "ms_per_step": best.ms_per_step,
            "default_ms_per_step": default[0].ms_per_step if default else None,
            "trials": [dataclasses.asdict(t) for t in trials]}


def signature(specs: Iterable) -> str:
    """Cache key of a workload: member count and total cells, rounded to two
    significant digits, so that similar mixes share a setting."""
    specs = list(specs)
    cells = sum(s.cells for s in specs)
    return "n%d_cells%.1e" % (len(specs), float("%.2g" % cells))


class TuneCache:
    """Settings by workload signature, kept in a JSON file."""

    def __init__(self, path: str):
        self.path = path
        try:
            with open(path) as f:
                self.table: Dict[str, dict] = json.load(f)
        except (OSError, ValueError):
            self.table = {}

    def get(self, specs) -> Optional[dict]:
        return self.table.get(signature(specs))

    def put(self, specs, result: dict) -> None:
        self.table[signature(specs)] = {k: v for k, v in result.items() if k != "trials"}
        with open(self.path, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)

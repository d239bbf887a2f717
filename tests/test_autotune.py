"""The batch autotuner's search on the host: with a stand-in batch whose
step cost depends on the launch shape and the colouring block (a fake clock
advanced by run_steady), tune_launch and tune pick the cheapest setting,
leave it set on the batch, time every candidate, and the cache keeps the
choice by workload signature."""
import os

from rocketcombustor_amd import autotune, nozzle


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeCase:
    def __init__(self, spec, block):
        self.spec, self.block, self.closed = spec, block, False

    def close(self):
        self.closed = True


class FakeBatch:
    """Step cost (seconds) = f(shape, block); run_steady advances the clock."""

    def __init__(self, clock, cost):
        self.clock, self.cost = clock, cost
        self.members, self.launch, self.calls = [], (256, 64, -1), []

    def add(self, g):
        self.members.append(g)

    def remove(self, g):
        self.members.remove(g)

    def info(self):
        return {"members": len(self.members), "cells": 0, "levels": 5}

    def set_launch(self, c, r, x):
        self.launch = (c, r, x)
        self.calls.append(self.launch)

    def run_steady(self, n, every=0, history=True):
        block = self.members[0].block if self.members else 0
        self.clock.t += n * self.cost(self.launch, block)

    def close(self):
        pass


def cost(shape, block):
    c, r, x = shape
    # cheapest: 128 cell threads, 128 run threads, XCD banding on, block 4
    return 1e-3 * (1.0 + abs(c - 128) / 256 + abs(r - 128) / 512 + (0.1 if x == 0 else 0.0) + abs(block - 4) / 8)


def test_tune_launch_picks_the_cheapest_shape_and_sets_it():
    clock = Clock()
    b = FakeBatch(clock, cost)
    b.add(FakeCase(None, 4))
    best, trials = autotune.tune_launch(b, steps=5, warmup=1, timer=clock, block=4)
    assert len(trials) == len(autotune.SHAPES) == 18
    assert (best.cell_threads, best.run_threads, best.xcd) == (128, 128, 1)
    assert b.launch == (128, 128, 1)
    for t in trials:   # every trial timed its own shape
        assert abs(t.ms_per_step - 1e3 * cost((t.cell_threads, t.run_threads, t.xcd), 4)) < 1e-9


def test_tune_over_blocks_and_cache(tmp_path):
    clock = Clock()
    specs = nozzle.sample_cases(3, cells_lo=800, cells_hi=3000)
    made = []

    def make_case(spec, block):
        made.append(FakeCase(spec, block))
        return made[-1]

    res = autotune.tune(specs, blocks=(2, 4, 8), steps=3, warmup=1, make_case=make_case,
                        make_batch=lambda: FakeBatch(clock, cost), timer=clock)
    assert (res["block"], res["cell_threads"], res["run_threads"], res["xcd"]) == (4, 128, 128, 1)
    assert len(res["trials"]) == 3 * 18
    assert abs(res["default_ms_per_step"] - 1e3 * cost((256, 64, 1), 4)) < 1e-9
    assert res["ms_per_step"] < res["default_ms_per_step"]
    assert all(g.closed for g in made) and len(made) == 9
    path = os.path.join(str(tmp_path), "tune.json")
    cache = autotune.TuneCache(path)
    assert cache.get(specs) is None
    cache.put(specs, res)
    again = autotune.TuneCache(path).get(specs)
    assert again["block"] == 4 and "trials" not in again
    assert autotune.signature(specs) == autotune.signature(list(specs))

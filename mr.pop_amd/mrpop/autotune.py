"""Autotuner of the batched form.

The batched engine runs many problems per launch; its throughput depends on
how much work each round creates: expansions per round (num_new_configs),
roadmap edges checked per new roadmap node (neighbour_candidates), dRRT
rounds queued per host step (rounds_per_step) and the check tiling of a
batch of several arm counts (arm_split: one mixed-arm launch with tiles of
320/A samples, or one launch per arm count with tiles of 64 samples).  The tuner runs short batches
of a calibration set for each candidate, discards candidates that leave a
problem unsolved or raise the median cost by more than `cost_slack` over the
best candidate, and returns the fastest of the rest.  Granularity is not
tuned: it sets the collision-check resolution and belongs to the problem.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

DEFAULT_GRID = {
    "num_new_configs": (128, 256, 512),
    "neighbour_candidates": (16, 32),
    "rounds_per_step": (8,),
    "arm_split": (0, 1),
}


@dataclass
class Trial:
    params: dict
    wall_s: float
    solved: int
    problems: int
    median_cost: float

    @property
    def problems_per_s(self) -> float:
        return self.problems / self.wall_s


@dataclass
class TuneResult:
    best: dict
    trials: list = field(default_factory=list)
    seconds: float = 0.0

    def as_dict(self) -> dict:
        return {"best": self.best, "seconds": self.seconds,
                "trials": [dict(t.params, problems_per_s=t.problems_per_s, solved=t.solved, problems=t.problems,
                                median_cost=t.median_cost) for t in self.trials]}


def tune(ctx, problems, base_settings: dict, base_options: dict, grid: dict | None = None,
         cost_slack: float = 0.02) -> TuneResult:
    """Pick settings for solve_batch on problems like `problems`.

    base_settings / base_options: keyword arguments of _lib.settings /
    _lib.solve_options_default shared by every candidate.
    """
    grid = grid or DEFAULT_GRID
    keys = list(grid)
    t_start = time.perf_counter()
    trials = []
    for values in itertools.product(*(grid[k] for k in keys)):
        params = dict(zip(keys, values))
        st_kw = dict(base_settings)
        op_kw = dict(base_options)
        for k, v in params.items():
            (st_kw if k == "num_new_configs" else op_kw)[k] = v
        st = _lib.settings(**st_kw)
        op = _lib.solve_options_default(**op_kw)
        t0 = time.perf_counter()
        sols = _lib.solve_batch(ctx, problems, [st] * len(problems), [op] * len(problems))
        wall = time.perf_counter() - t0
        costs = [s.cost for s in sols if s.solved]
        trials.append(Trial(params, wall, len(costs), len(problems), float(np.median(costs)) if costs else np.inf))
    full = [t for t in trials if t.solved == t.problems] or trials
    best_cost = min(t.median_cost for t in full)
    ok = [t for t in full if t.median_cost <= best_cost * (1 + cost_slack)]
    best = max(ok, key=lambda t: t.problems_per_s)
    return TuneResult(best=dict(best.params), trials=trials, seconds=time.perf_counter() - t_start)

"""Autotuner of the batched form.

The batched engine runs many problems per launch; its throughput depends on
how much work each round creates and how it is laid out on the GPU:
expansions per round (num_new_configs), samples per checked edge
(granularity), roadmap edges checked per new roadmap node
(neighbour_candidates) and the check tiling of a batch of several arm counts
(arm_split: one mixed-arm launch with tiles of 320/A samples per block, or
one launch per arm count with tiles of 64 samples).

The search runs on a calibration set of problems of other seeds, as many as
the batches it tunes for: throughput depends on the load of the whole
batch, and a few problems leave the GPU latency bound, where more work per
problem looks cheap.  Calibration problems the starting candidate (the
smallest num_new_configs at granularity 32) leaves unsolved within the
budget are dropped first.  It is a coordinate search over the grid:
  1. num_new_configs in increasing order at granularity 32, stopping once
     throughput has fallen twice in a row (the work of a round grows with
     it, its parallelism saturates);
  2. the other granularities at the best num_new_configs;
  3. the tiling and neighbour-candidate choices at the best pair.
A candidate is valid when it solves every calibration problem and every
segment of every returned path is collision-free when re-checked by the
device checker at the problems' own granularity (that of the base settings),
at the chunking the planner commits segments with; among valid candidates
whose median cost is within `cost_slack` of the best, the fastest wins.  A
coarser granularity is therefore only chosen when its paths hold up at the
resolution the problems ask for.
"""
from __future__ import annotations

import itertools
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

GRID = {
    "num_new_configs": (256, 512, 1024, 2048, 4096),
    "granularity": (16, 32, 64),
    "neighbour_candidates": (16, 32),
    "arm_split": (0, 1),
}
START = {"granularity": 32, "neighbour_candidates": 16, "arm_split": 0}
SETTING_KEYS = ("num_new_configs", "granularity", "max_iters")
VALIDATION_GRANULARITY = 64   # when the base settings name no granularity


@dataclass
class Trial:
    params: dict
    wall_s: float      # until 90 % of the problems were done
    solved: int
    problems: int
    median_cost: float
    valid: bool = True
    stage: int = 1

    @property
    def problems_per_s(self) -> float:
        """Throughput of the bulk of the batch: 90 % of its problems over the
        time they took (one straggler would otherwise set the figure of every
        candidate alike)."""
        return 0.9 * self.problems / self.wall_s


@dataclass
class TuneResult:
    best: dict
    trials: list = field(default_factory=list)
    seconds: float = 0.0
    excluded: int = 0   # calibration problems the starting candidate left unsolved (not used further)

    def as_dict(self) -> dict:
        return {"best": self.best, "seconds": self.seconds, "excluded_unsolved": self.excluded,
                "trials": [dict(t.params, problems_per_s=t.problems_per_s, solved=t.solved, problems=t.problems,
                                median_cost=t.median_cost, paths_valid=t.valid, stage=t.stage)
                           for t in self.trials]}


def path_free(world, path, rng, granularity=VALIDATION_GRANULARITY) -> bool:
    """Every segment of `path` free at `granularity` samples per chunk, for
    one of the chunkings the planner commits segments with: floor(d/range)+1
    chunks (connections, pop_planner.cu:657-690), ceil(d/range) chunks
    (simplified segments, path_simplification.cu:159-162) or one edge (tree
    extensions)."""
    path = np.asarray(path, np.float64)
    for a, b in zip(path[:-1], path[1:]):
        d = float(np.linalg.norm(b - a))
        ok = False
        for k in sorted({1, max(1, math.ceil(d / rng)), int(d / rng) + 1}):
            frm = np.array([a + (b - a) * c / k for c in range(k)], np.float32)
            to = np.array([a + (b - a) * (c + 1) / k for c in range(k)], np.float32)
            if not world.check_edges(_lib.MODE_FULL, granularity, frm, to).any():
                ok = True
                break
        if not ok:
            return False
    return True


def _run(ctx, problems, params, base_settings, base_options, worlds, stage, keep_sols=None):
    st_kw = dict(base_settings)
    op_kw = dict(base_options)
    for k, v in params.items():
        (st_kw if k in SETTING_KEYS else op_kw)[k] = v
    st = _lib.settings(**st_kw)
    op = _lib.solve_options_default(**op_kw)
    t0 = time.perf_counter()
    sols = _lib.solve_batch(ctx, problems, [st] * len(problems), [op] * len(problems))
    wall = time.perf_counter() - t0
    done = [s.info["wall_ns"] * 1e-9 for s in sols if getattr(s, "info", None)]
    if done:
        wall = float(np.percentile(done, 90))
    if keep_sols is not None:
        keep_sols.extend(sols)
    costs = [s.cost for s in sols if s.solved]
    rng = float(st_kw.get("range", 2.0))
    g = int(base_settings.get("granularity", VALIDATION_GRANULARITY))
    valid = all(path_free(worlds[id(p)], s.path, rng, g) for p, s in zip(problems, sols) if s.solved) \
        if worlds is not None else True
    return Trial(dict(params), wall, len(costs), len(problems), float(np.median(costs)) if costs else np.inf,
                 valid, stage)


def _ranked(trials, cost_slack):
    ok = [t for t in trials if t.solved == t.problems and t.valid] or [t for t in trials if t.valid] or trials
    best_cost = min(t.median_cost for t in ok)
    good = [t for t in ok if t.median_cost <= best_cost * (1 + cost_slack)]
    return sorted(good, key=lambda t: -t.problems_per_s)


def tune(ctx, problems, base_settings: dict, base_options: dict, grid: dict | None = None,
         cost_slack: float = 0.02, validate: bool = True, log=None) -> TuneResult:
    """Pick settings for solve_batch on problems like `problems`.

    base_settings / base_options: keyword arguments of _lib.settings /
    _lib.solve_options_default shared by every candidate (the tuned keys
    override them; a time_budget_s there bounds each calibration problem, and
    a problem it cuts short counts as unsolved).  validate: re-check returned
    paths at the base settings' granularity.  log: called with a line per
    trial."""
    grid = grid or GRID
    t_start = time.perf_counter()
    worlds = {id(p): ctx.world(p.scene, p.n_arms, p.bases) for p in problems} if validate else None
    trials = []
    # 0. the starting candidate; calibration problems it leaves unsolved
    # within the budget are dropped: one such problem holds every candidate's
    # batch until the budget ends and hides the differences between them
    cur = dict(START, num_new_configs=grid["num_new_configs"][0])
    cur["granularity"] = START["granularity"] if START["granularity"] in grid["granularity"] else grid["granularity"][0]
    first = []
    trials.append(_run(ctx, problems, cur, base_settings, base_options, worlds, 0, first))
    kept = [p for p, s in zip(problems, first) if s.solved] or problems
    excluded = len(problems) - len(kept)
    problems = kept
    if log:
        log(f"{cur}: {trials[0].problems_per_s:.2f} problems/s, {trials[0].solved}/{trials[0].problems} solved; "
            f"{excluded} unsolved problems leave the calibration set")

    def run(params, stage):
        t = _run(ctx, problems, params, base_settings, base_options, worlds, stage)
        trials.append(t)
        if log:
            log(f"{params}: {t.problems_per_s:.2f} problems/s, {t.solved}/{t.problems} solved, "
                f"median cost {t.median_cost:.3f}, paths valid {t.valid}")
        return t

    # 1. num_new_configs, ascending, until throughput falls twice in a row
    falls, last = 0, None
    for nnc in grid["num_new_configs"]:
        t = run(dict(cur, num_new_configs=nnc), 1)
        if last is not None and t.problems_per_s < last:
            falls += 1
            if falls == 2:
                break
        else:
            falls = 0
        last = t.problems_per_s
    best = _ranked([t for t in trials if t.stage >= 1], cost_slack)[0]
    # 2. the other granularities at that num_new_configs
    for g in grid["granularity"]:
        if g != best.params["granularity"]:
            run(dict(best.params, granularity=g), 2)
    best = _ranked([t for t in trials if t.stage >= 1], cost_slack)[0]
    # 3. tiling and roadmap candidates at that pair
    for kc, split in itertools.product(grid["neighbour_candidates"], grid["arm_split"]):
        params = dict(best.params, neighbour_candidates=kc, arm_split=split)
        if params != best.params:
            run(params, 3)
    best = _ranked([t for t in trials if t.stage >= 1], cost_slack)[0]
    for w in (worlds or {}).values():
        w.close()
    return TuneResult(best=dict(best.params), trials=trials, seconds=time.perf_counter() - t_start,
                      excluded=excluded)

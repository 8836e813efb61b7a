"""Selection rules of the batched-form autotuner (mrpop/autotune.py), with
solve_batch and the path re-check replaced by stand-ins: no GPU."""
import types

import numpy as np

from mrpop import _lib, autotune


def _fake_batch(table, default):
    """solve_batch stand-in: per (num_new_configs, granularity,
    neighbour_candidates, arm_split), (seconds, solved, cost, path_ok)."""
    def solve_batch(ctx, problems, settings, options):
        key = (settings[0].num_new_configs, settings[0].granularity, options[0].neighbour_candidates,
               options[0].arm_split)
        secs, solved, cost, ok = table.get(key, default)
        solve_batch.calls.append(key)
        solve_batch.elapsed += secs
        return [types.SimpleNamespace(solved=i < solved, cost=cost, path=np.zeros((2, 7)) + (0 if ok else 1))
                for i in range(len(problems))]
    solve_batch.elapsed = 0.0
    solve_batch.calls = []
    return solve_batch


class _Ctx:
    def world(self, *a):
        return types.SimpleNamespace(close=lambda: None)


def _run(monkeypatch, table, default=(10.0, 4, 10.0, True), slack=0.02):
    fake = _fake_batch(table, default)
    monkeypatch.setattr(_lib, "solve_batch", fake)
    # a path re-check stand-in: the fake paths carry their verdict
    monkeypatch.setattr(autotune, "path_free", lambda world, path, rng, g=64: not path.any())
    # the tuner's clock advances only inside solve_batch, by the table's time
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: fake.elapsed)
    probs = [types.SimpleNamespace(scene=None, n_arms=2, bases=None) for _ in range(4)]
    return autotune.tune(_Ctx(), probs, {}, {}, cost_slack=slack), fake.calls


def test_grid_covers_the_batched_configuration():
    """num_new_configs 256..4096 and granularity 16/32/64 (the batched
    benchmark configuration), the check tiling and roadmap candidates."""
    assert autotune.GRID["num_new_configs"] == (256, 512, 1024, 2048, 4096)
    assert autotune.GRID["granularity"] == (16, 32, 64)
    for k in ("neighbour_candidates", "arm_split"):
        assert hasattr(_lib.solve_options_default(), k)
    for k in autotune.SETTING_KEYS:
        assert hasattr(_lib.settings_default(), k)


def test_coordinate_search_and_validity(monkeypatch):
    table = {(256, 32, 16, 0): (1.0, 4, 10.0, True),
             (512, 32, 16, 0): (0.8, 4, 10.0, True),
             (1024, 32, 16, 0): (0.9, 4, 10.0, True),    # slower: one fall
             (2048, 32, 16, 0): (1.5, 4, 10.0, True),    # slower again: the scan stops before 4096
             (512, 16, 16, 0): (0.5, 4, 10.0, False),    # fastest, but its paths fail the re-check at the problems' granularity
             (512, 64, 16, 0): (1.2, 4, 10.0, True),
             (512, 32, 32, 1): (0.7, 4, 10.0, True)}
    res, calls = _run(monkeypatch, table)
    assert (4096, 32, 16, 0) not in calls
    assert res.best == {"num_new_configs": 512, "granularity": 32, "neighbour_candidates": 32, "arm_split": 1}
    bad = [t for t in res.trials if t.params["granularity"] == 16][0]
    assert not bad.valid
    assert len(calls) == 1 + 4 + 2 + 3  # the starting candidate runs first on the whole set


def test_problems_the_start_leaves_unsolved_leave_the_calibration_set(monkeypatch):
    table = {(256, 32, 16, 0): (1.0, 3, 10.0, True)}   # the starting candidate solves 3 of 4
    res, _ = _run(monkeypatch, table)
    assert res.excluded == 1
    assert [t.problems for t in res.trials][:2] == [4, 3]


def test_unsolved_and_costly_candidates_lose(monkeypatch):
    table = {(256, 32, 16, 0): (1.0, 4, 10.0, True),
             (512, 32, 16, 0): (0.5, 3, 10.0, True),     # fast, leaves a problem unsolved
             (1024, 32, 16, 0): (0.6, 4, 11.0, True),    # fast, 10 % worse median cost
             (2048, 32, 16, 0): (0.9, 4, 10.1, True)}    # within 2 % of the best cost
    res, _ = _run(monkeypatch, table)
    assert res.best["num_new_configs"] == 2048
    res, _ = _run(monkeypatch, table, slack=0.2)
    assert res.best["num_new_configs"] == 1024

"""Selection rules of the batched-form autotuner (mrpop/autotune.py), with
solve_batch replaced by a stand-in: no GPU."""
import types

from mrpop import _lib, autotune


def _fake_batch(table):
    """solve_batch stand-in: per num_new_configs, (seconds, solved, cost)."""
    def solve_batch(ctx, problems, settings, options):
        secs, solved, cost = table[(settings[0].num_new_configs, options[0].neighbour_candidates)]
        solve_batch.elapsed += secs
        return [types.SimpleNamespace(solved=i < solved, cost=cost) for i in range(len(problems))]
    solve_batch.elapsed = 0.0
    return solve_batch


def _run(monkeypatch, table, grid, slack=0.02):
    fake = _fake_batch(table)
    monkeypatch.setattr(_lib, "solve_batch", fake)
    # the tuner's clock advances only inside solve_batch, by the table's time
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: fake.elapsed)
    return autotune.tune(None, [object()] * 4, {}, {}, grid=grid, cost_slack=slack)


def test_picks_the_fastest_candidate_that_solves_everything(monkeypatch):
    grid = {"num_new_configs": (128, 256, 512), "neighbour_candidates": (16,)}
    table = {(128, 16): (0.5, 3, 10.0),   # fastest, but leaves a problem unsolved
             (256, 16): (1.0, 4, 10.0),
             (512, 16): (2.0, 4, 10.0)}
    res = _run(monkeypatch, table, grid)
    assert res.best == {"num_new_configs": 256, "neighbour_candidates": 16}
    assert len(res.trials) == 3


def test_cost_slack_excludes_faster_but_worse_paths(monkeypatch):
    grid = {"num_new_configs": (128, 256), "neighbour_candidates": (16, 32)}
    table = {(128, 16): (0.5, 4, 11.0),   # 10 % worse median cost than the best
             (128, 32): (0.6, 4, 10.1),   # within 2 % of the best
             (256, 16): (1.0, 4, 10.0),
             (256, 32): (1.2, 4, 10.0)}
    res = _run(monkeypatch, table, grid)
    assert res.best == {"num_new_configs": 128, "neighbour_candidates": 32}
    res = _run(monkeypatch, table, grid, slack=0.2)
    assert res.best == {"num_new_configs": 128, "neighbour_candidates": 16}


def test_default_grid_covers_the_tiling():
    assert set(autotune.DEFAULT_GRID) == {"num_new_configs", "neighbour_candidates", "rounds_per_step", "arm_split"}
    for k in autotune.DEFAULT_GRID:
        if k != "num_new_configs":
            assert hasattr(_lib.solve_options_default(), k)

"""Planner parity on the GPU: dRRT + AO-x through the C ABI.

A returned path must (a) run from a goal to the start, (b) stay inside the
joint limits, (c) cost the sum of its segment lengths, and (d) be collision
free at the resolution the planner checks it with.  (d) is decided by the
FP64 oracle: a segment is free if, for one of the discretisations the
planner uses for the edges it commits, every chunk is free with contacts
closer than 1e-5 m read as free:

* a tree extension edge: one edge of ``granularity`` samples
  (pop_planner.cu:521-546 checks it with the same ``granularity``),
* a tree connection: ``floor(d / range) + 1`` chunks (pop_planner.cu:657-690),
* a simplified segment: ``ceil(d / range)`` chunks
  (path_simplification.cu:159-162).

The reference's planner output is random (curand streams and scheduling), so
its own tests check properties like these rather than paths; here the
planner is also deterministic for a seed, and a problem gives the same path
whether it is solved alone or in a batch with others.
"""
import dataclasses
import math
import time

import numpy as np
import pytest

from mrpop import _lib, geometry, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# settings of the fixture runs: 300 new configurations per round, 64 samples
# per edge, range 2.0, three AO-x passes, short optimisation passes
SETTINGS = dict(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=3, optimize_iters=1500)
BUDGET_S = 60.0


def _settings(**kw):
    d = dict(SETTINGS)
    d.update(kw)
    return _lib.settings(**d)


def _options(**kw):
    d = dict(time_budget_s=BUDGET_S)
    d.update(kw)
    return _lib.solve_options_default(**d)


def _fixture_problems():
    four = scenes.load_problem_file("tests/golden/panda_four_problems.json", geometry.default_bases(4))
    five = scenes.load_problem_file("tests/golden/panda_five_problems.json", geometry.default_bases(5))
    return four[:3] + five[:2]


PROBLEMS = _fixture_problems()
NAMES = [p.name for p in PROBLEMS]


def _chunks(a, b, n):
    return [(a + (b - a) * c / n, a + (b - a) * (c + 1) / n) for c in range(n)]


def segment_verdicts(problem, path, granularity, rng):
    """Per segment: True if free under one of the planner's discretisations."""
    path = np.asarray(path, np.float64)
    rows, owner = [], []
    for i in range(len(path) - 1):
        a, b = path[i], path[i + 1]
        d = float(np.linalg.norm(b - a))
        for variant, n in enumerate((1, max(1, math.ceil(d / rng)), int(d / rng) + 1)):
            for f, t in _chunks(a, b, n):
                rows.append((f, t))
                owner.append((i, variant))
    if not rows:
        return []
    frm = np.array([r[0] for r in rows], np.float32)
    to = np.array([r[1] for r in rows], np.float32)
    r = O.check_edges(problem.scene, O.MODE_FULL, problem.bases, granularity, frm, to)
    hit = {}
    for (i, v), lo in zip(owner, r["lo"]):
        hit[(i, v)] = hit.get((i, v), False) or bool(lo)
    return [any(not hit[(i, v)] for v in range(3)) for i in range(len(path) - 1)]


def interior_waypoint(path):
    """The middle waypoint of a path among those equal to neither end."""
    inner = [w for w in path if not (np.array_equal(w, path[0]) or np.array_equal(w, path[-1]))]
    if not inner:
        pytest.skip("the path has no interior waypoint")
    return inner[len(inner) // 2]


def check_solution(problem, sol, granularity=SETTINGS["granularity"], rng=SETTINGS["range"]):
    assert sol.solved, f"{problem.name}: not solved ({sol.info})"
    path = sol.path
    assert path.shape[1] == problem.dim and len(path) >= 2
    assert np.allclose(path[-1], problem.start, atol=1e-6), "path must end at the start"
    assert min(np.abs(path[0] - g).max() for g in problem.goals) < 1e-6, "path must begin at a goal"
    low = np.tile(np.array(geometry.table("kLow"), np.float32), problem.n_arms)
    high = low + np.tile(np.array(geometry.table("kScale"), np.float32), problem.n_arms)
    assert (path >= low - 1e-5).all() and (path <= high + 1e-5).all(), "path leaves the joint limits"
    length = float(np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1).sum())
    assert sol.cost == pytest.approx(length, rel=1e-3, abs=1e-3)
    free = segment_verdicts(problem, path, granularity, rng)
    assert all(free), f"{problem.name}: segments in collision: {[i for i, f in enumerate(free) if not f]}"


@pytest.fixture(scope="module")
def alone(ctx):
    out = {}
    for p in PROBLEMS:
        print(f"[plan] solve {p.name}", flush=True)
        out[p.name] = _lib.solve(ctx, p, _settings(), _options())
    return out


@pytest.mark.parametrize("name", NAMES)
def test_fixture_problem_solved_and_free(alone, name):
    p = PROBLEMS[NAMES.index(name)]
    sol = alone[name]
    assert sol.info["status"] == _lib.OK, sol.info
    check_solution(p, sol)


@pytest.mark.parametrize("name", NAMES)
def test_anytime_improvements(alone, name):
    """Every recorded solution improves on the one before and is itself a valid path."""
    p = PROBLEMS[NAMES.index(name)]
    sol = alone[name]
    costs = [c for c, _ in sol.seed_paths]
    assert len(costs) >= 1 and len(costs) == len(sol.improve_s)
    assert all(b < a for a, b in zip(costs, costs[1:])), costs
    assert sol.cost == pytest.approx(costs[-1])
    assert sol.improve_s == sorted(sol.improve_s)
    assert sol.ttfs_s == pytest.approx(sol.improve_s[0]) and 0 < sol.ttfs_s <= sol.wall_s
    for c, path in sol.seed_paths:
        assert len(path) >= 2
        assert np.allclose(path[-1], p.start, atol=1e-6)


def test_deterministic_for_a_seed(ctx, alone):
    p = PROBLEMS[0]
    again = _lib.solve(ctx, p, _settings(), _options())
    assert np.array_equal(again.path, alone[p.name].path)
    assert again.info["rounds"] == alone[p.name].info["rounds"]


def test_seed_changes_the_search(ctx, alone):
    p = PROBLEMS[1]
    q = dataclasses.replace(p, seed=p.seed + 1)
    other = _lib.solve(ctx, q, _settings(), _options())
    check_solution(q, other)
    assert not np.array_equal(other.path, alone[p.name].path)


def test_same_alone_and_in_the_batch(ctx, alone):
    """Slots share every launch but not their state: batching changes no result."""
    sols = _lib.solve_batch(ctx, PROBLEMS, [_settings()] * len(PROBLEMS), [_options()] * len(PROBLEMS))
    for p, s in zip(PROBLEMS, sols):
        assert s.info["status"] == _lib.OK
        assert np.array_equal(s.path, alone[p.name].path), p.name
        assert [c for c, _ in s.seed_paths] == [c for c, _ in alone[p.name].seed_paths]
        assert s.info["rounds"] == alone[p.name].info["rounds"]


def test_probe_pass_changes_no_result(ctx, alone):
    """Checking planner edges at a few samples first, and in full only while
    free, gives the verdicts of one full pass: every result is unchanged."""
    sols = _lib.solve_batch(ctx, PROBLEMS, [_settings()] * len(PROBLEMS),
                            [_options(probe_samples=4)] * len(PROBLEMS))
    for p, s in zip(PROBLEMS, sols):
        assert s.info["status"] == _lib.OK
        assert np.array_equal(s.path, alone[p.name].path), p.name
        assert [c for c, _ in s.seed_paths] == [c for c, _ in alone[p.name].seed_paths]
        assert s.info["rounds"] == alone[p.name].info["rounds"]
    mixed = [scenes.make_problem(k, 3, _lib.gpu_checker(ctx)) for k in ("plain2", "table3", "bin4", "shelf5")]
    one = _lib.solve_batch(ctx, mixed, [_settings()] * 4, [_options()] * 4)
    two = _lib.solve_batch(ctx, mixed, [_settings()] * 4, [_options(probe_samples=4)] * 4)
    for a, b in zip(one, two):
        assert np.array_equal(a.path, b.path) and a.info["rounds"] == b.info["rounds"]
    # the other tiling: one check launch per arm count instead of the mixed kernel
    three = _lib.solve_batch(ctx, mixed, [_settings()] * 4, [_options(arm_split=1)] * 4)
    for a, b in zip(one, three):
        assert np.array_equal(a.path, b.path) and a.info["rounds"] == b.info["rounds"]


def test_graph_replay_changes_no_result(ctx, alone):
    """dRRT rounds replayed from captured HIP graphs and launched one kernel
    at a time (the default) give the same results; a solve replays each
    captured sequence many times, with other lanes' work between."""
    replayed = [_lib.solve(ctx, p, _settings(), _options(graph_rounds=1)) for p in PROBLEMS[:3]]
    for p, s in zip(PROBLEMS[:3], replayed):
        assert np.array_equal(s.path, alone[p.name].path), p.name
        assert s.info["rounds"] == alone[p.name].info["rounds"]
    mixed = [scenes.make_problem(k, 5, _lib.gpu_checker(ctx)) for k in ("plain2", "table3", "bin4", "shelf5")]
    one = _lib.solve_batch(ctx, mixed, [_settings()] * 4, [_options()] * 4)
    two = _lib.solve_batch(ctx, mixed, [_settings()] * 4, [_options(graph_rounds=1)] * 4)
    for a, b in zip(one, two):
        assert np.array_equal(a.path, b.path) and a.info["rounds"] == b.info["rounds"]
        assert [c for c, _ in a.seed_paths] == [c for c, _ in b.seed_paths]


def test_roadmap_reassign(ctx):
    """roadmap_reassign: arms whose start and goal are connected stop
    sampling, the others go on; paths stay valid and the connected arms'
    roadmaps stop growing."""
    probs = [scenes.make_problem(k, 7, _lib.gpu_checker(ctx)) for k in ("bin4", "shelf5")]
    for p in probs:
        base = _lib.solve(ctx, p, _settings(), _options())
        re = _lib.solve(ctx, p, _settings(roadmap_reassign=1), _options())
        check_solution(p, re)
        assert sum(re.info["roadmap_size"]) <= sum(base.info["roadmap_size"]), p.name
        print(f"[reassign] {p.name}: roadmap {base.info['roadmap_size']} -> {re.info['roadmap_size']}, "
              f"ttfs {base.ttfs_s:.4f} -> {re.ttfs_s:.4f} s", flush=True)


@pytest.mark.parametrize("kind,seed", [("plain2", 3), ("table3", 5)])
def test_generated_two_and_three_arm(ctx, kind, seed):
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    sol = _lib.solve(ctx, p, _settings(), _options())
    check_solution(p, sol)


def test_without_simplification(ctx):
    p = PROBLEMS[0]
    sol = _lib.solve(ctx, p, _settings(path_simplify=False, rrtc_iter=1), _options())
    check_solution(p, sol)
    assert sol.info["passes"] == 1


def test_server_continuous_batching(ctx, alone):
    """Problems submitted while others run finish with their solo results."""
    srv = _lib.Server(ctx)
    tickets = {}
    tickets[srv.submit(PROBLEMS[0], _settings(), _options())] = PROBLEMS[0]
    steps = 0
    pending = list(PROBLEMS[1:4])
    cancelled = srv.submit(PROBLEMS[4], _settings(), _options())
    srv.cancel(cancelled)
    assert srv.poll(cancelled)[0] == 2
    deadline = time.time() + 3 * BUDGET_S
    while True:
        if pending and steps % 50 == 0:
            p = pending.pop(0)
            tickets[srv.submit(p, _settings(), _options())] = p
        active, waiting = srv.step()
        steps += 1
        if steps == 10:  # anytime: a snapshot while running is a valid (maybe unsolved) result
            st, snap = srv.poll(next(iter(tickets)), snapshot=True)
            assert st in (1, 2) and snap is not None
        if not pending and active == 0 and waiting == 0:
            break
        assert time.time() < deadline, "server did not finish"
    assert srv.bytes() == 0
    for t, p in tickets.items():
        st, sol = srv.poll(t, snapshot=True)
        assert st == 2
        assert np.array_equal(sol.path, alone[p.name].path), p.name
    assert srv.poll(123456789)[0] == -1
    srv.close()


def test_server_budget(ctx):
    """A problem larger than the whole budget is refused, not queued forever:
    it comes back done with status ERR_BUDGET and the requests behind it are
    admitted and planned.  A batch with one such problem plans the others."""
    small = _options(tree_capacity=65536, roadmap_capacity=20000, neighbour_capacity=64)
    big = _options()
    assert _lib.slot_bytes(ctx, PROBLEMS[0].n_arms, _settings(), big) > \
        _lib.slot_bytes(ctx, PROBLEMS[1].n_arms, _settings(), small) + (1 << 20)
    srv = _lib.Server(ctx, hbm_budget_bytes=_lib.slot_bytes(ctx, PROBLEMS[1].n_arms, _settings(), small) + (1 << 20))
    t_big = srv.submit(PROBLEMS[0], _settings(), big)
    t_small = srv.submit(PROBLEMS[1], _settings(), small)
    deadline = time.time() + 120
    while srv.step() != (0, 0):
        assert time.time() < deadline, "server did not finish"
    st, sol = srv.poll(t_big, snapshot=True)
    assert st == 2 and sol.info["status"] == _lib.ERR_BUDGET and not sol.solved
    st, sol = srv.poll(t_small, snapshot=True)
    assert st == 2 and sol.info["status"] in (_lib.OK, _lib.ERR_TIMEOUT) and sol.solved
    check_solution(PROBLEMS[1], sol)
    srv.close()


def test_time_budget(ctx):
    p = PROBLEMS[3]
    sol = _lib.solve(ctx, p, _settings(optimize_iters=400000), _options(time_budget_s=1.0))
    assert sol.info["status"] == _lib.ERR_TIMEOUT
    assert sol.wall_s < 2.0
    if sol.solved:  # an anytime result survives the cut
        check_solution(p, sol)


def test_invalid_problems(ctx):
    p = PROBLEMS[0]
    bad = dataclasses.replace(p, n_arms=6, bases=np.zeros((6, 12), np.float32), start=np.zeros(42, np.float32),
                              goals=np.zeros((1, 42), np.float32))
    with pytest.raises(_lib.MrpopError):
        _lib.solve(ctx, bad)
    nog = dataclasses.replace(p, goals=p.goals[:0])
    with pytest.raises(_lib.MrpopError):
        _lib.solve(ctx, nog)


def test_invalid_problems_are_rejected(ctx):
    """Malformed problems fail the call; problems whose settings or values
    are rejected on admission come back unplanned with status ERR_INVALID
    while the rest of the batch is planned as usual."""
    p = PROBLEMS[0]
    with pytest.raises(_lib.MrpopError) as e:
        _lib.solve(ctx, dataclasses.replace(p, n_arms=1, start=p.start[:7], goals=p.goals[:, :7], bases=p.bases[:1]))
    assert e.value.code == _lib.ERR_INVALID
    nan_start = p.start.copy()
    nan_start[3] = np.nan
    inf_goal = p.goals.copy()
    inf_goal[0, 9] = np.inf
    ok = (_settings(), _options())
    # rejected before any size is computed from them: negative counts would
    # otherwise read as a slot larger than the device
    bad = [(p, _settings(granularity=0), ok[1]), (p, _settings(granularity=65), ok[1]),
           (p, _settings(num_new_configs=0), ok[1]), (p, _settings(num_new_configs=-5), ok[1]),
           (p, _settings(range=0.0), ok[1]), (p, _settings(range=float("nan")), ok[1]),
           (p, ok[0], _options(tree_capacity=-1)), (p, ok[0], _options(neighbour_candidates=65)),
           # the goal tree's roots would not fit: refused on the host, no kernel runs
           (p, ok[0], _options(tree_capacity=1)), (p, _settings(max_samples=1), ok[1]),
           (dataclasses.replace(p, goals=np.stack([p.goals[0]] * 3)), ok[0], _options(tree_capacity=3)),
           (dataclasses.replace(p, goals=np.stack([p.goals[0]] * 3)), _settings(max_samples=2), ok[1]),
           (dataclasses.replace(p, start=nan_start), *ok), (dataclasses.replace(p, goals=inf_goal), *ok),
           # steering period, and a roadmap too small to keep room for another query's nodes
           (p, ok[0], _options(biased_iter=-1)), (p, ok[0], _options(roadmap_cache=1, roadmap_capacity=35))]
    cases = [(p, *ok)] + bad + [(PROBLEMS[3], *ok)]
    probs = [c[0] for c in cases]
    sols = _lib.solve_batch(ctx, probs, [c[1] for c in cases], [c[2] for c in cases])
    for q, sol in zip(probs[1:-1], sols[1:-1]):
        assert sol.info["status"] == _lib.ERR_INVALID and not sol.solved and len(sol.path) == 0
    for q, sol in (probs[0], sols[0]), (probs[-1], sols[-1]):
        assert sol.info["status"] == _lib.OK and sol.solved
        check_solution(q, sol)
    # the server reports a rejected request as done, unplanned, and holds no memory for it
    srv = _lib.Server(ctx)
    t = srv.submit(p, _settings(num_new_configs=-5), _options())
    assert srv.step() == (0, 0)
    st, sol = srv.poll(t, snapshot=True)
    assert st == 2 and sol.info["status"] == _lib.ERR_INVALID and not sol.solved
    assert srv.bytes() == 0
    srv.close()


def test_several_goals(ctx):
    """Problems with several goals: every returned path starts at one of them
    (the one its goal-tree chain reached), costs its length and is free.  In
    the second set goal 1 is a waypoint halfway along a solved path, closer
    than goal 0, so paths end at a goal other than the first."""
    checker = _lib.gpu_checker(ctx)
    probs = [scenes.make_problem(kind, seed, checker, n_goals=3)
             for kind, seed in (("table3", 21), ("bin4", 22), ("shelf5", 23), ("plain2", 24))]
    first = _lib.solve_batch(ctx, probs, [_settings()] * len(probs), [_options()] * len(probs))
    near = []
    for p, sol in zip(probs, first):
        check_solution(p, sol)
        mid = interior_waypoint(sol.path)
        near.append(dataclasses.replace(p, goals=np.stack([p.goals[0], mid, p.goals[2]])))
    sols = _lib.solve_batch(ctx, near, [_settings()] * len(near), [_options()] * len(near))
    reached = []
    for p, sol in zip(near, sols):
        check_solution(p, sol)
        reached.append(int(np.argmin([np.abs(sol.path[0] - g).max() for g in p.goals])))
    print(f"[goals] goal reached per problem: {reached}", flush=True)
    assert any(r != 0 for r in reached)


# the reference's evaluation harness (scripts/evaluate_mr.cu:159-166): pop_settings
# defaults with num_new_configs 300, granularity 64, range 2.0, rrtc_iter 3, and
# dRRT::solve(..., n_arms, 8), i.e. block_per_robot 8
HARNESS = dict(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=3)
FIXTURE_BUDGET_S = 20.0


def test_reference_fixtures_all_solved(ctx):
    """The reference's own benchmark: all 50 bin-picking (4 arms) and 50
    shelf (5 arms) problems of scripts/panda_{four,five}_problems.json at the
    harness settings, planned together in one batch.  README.md:17-20 claims
    a 100% solve rate for both sets; every problem must be solved and every
    returned path must pass the FP64 re-validation.  The optimise passes run
    with the harness's optimize_iters (400 000) under a wall-clock budget
    per problem.  Solve rate, time to first solution and cost distribution
    are printed for the record."""
    four = scenes.load_problem_file("tests/golden/panda_four_problems.json", geometry.default_bases(4))
    five = scenes.load_problem_file("tests/golden/panda_five_problems.json", geometry.default_bases(5))
    probs = four + five
    assert len(four) == 50 and len(five) == 50
    st = _lib.settings(**HARNESS)
    op = _lib.solve_options_default(time_budget_s=FIXTURE_BUDGET_S)
    print(f"[fixtures] solving {len(probs)} problems in one batch", flush=True)
    t0 = time.time()
    sols = _lib.solve_batch(ctx, probs, [st] * len(probs), [op] * len(probs))
    wall = time.time() - t0
    unsolved = [p.name for p, s in zip(probs, sols) if not s.solved]
    for name, group in (("panda_four", sols[:50]), ("panda_five", sols[50:])):
        solved = [s for s in group if s.solved]
        costs = np.array([s.cost for s in solved])
        ttfs = np.array([s.ttfs_s for s in solved])
        first = np.array([s.info["n_seed_paths"] and s.seed_paths[0][0] for s in solved])
        print(f"[fixtures] {name}: solved {len(solved)}/50, cost median {np.median(costs):.3f} "
              f"p90 {np.percentile(costs, 90):.3f} (first solution median {np.median(first):.3f}), "
              f"TTFS median {np.median(ttfs) * 1e3:.0f} ms p99 {np.percentile(ttfs, 99) * 1e3:.0f} ms", flush=True)
    print(f"[fixtures] batch wall {wall:.1f} s", flush=True)
    assert not unsolved, f"unsolved: {unsolved}"
    for p, s in zip(probs, sols):
        assert s.info["union_find_bound"] == 0 and s.info["edge_table_full"] == 0
        check_solution(p, s, granularity=HARNESS["granularity"], rng=HARNESS["range"])


def test_stuck_step_ends_its_problems_with_timeout():
    """A planner step whose kernels outlive the context's lane deadline ends
    its problems with ERR_TIMEOUT instead of blocking the caller.  The
    condition is forced from the host: a deadline of 0 s while the kernels
    run and complete normally.  The context plans normally afterwards."""
    import ctypes
    lib = _lib.lib()
    lib.mrpop_diag_set_lane_deadline.argtypes = [ctypes.c_void_p, ctypes.c_double]
    c = _lib.Context(0)
    try:
        assert lib.mrpop_diag_set_lane_deadline(c.handle, 0.0) == _lib.OK
        t0 = time.perf_counter()
        sol = _lib.solve(c, PROBLEMS[0], _settings(), _options())
        assert sol.info["status"] == _lib.ERR_TIMEOUT and not sol.solved
        assert time.perf_counter() - t0 < 10.0
        assert lib.mrpop_diag_set_lane_deadline(c.handle, 120.0) == _lib.OK
        sol = _lib.solve(c, PROBLEMS[0], _settings(), _options())
        assert sol.info["status"] == _lib.OK
        check_solution(PROBLEMS[0], sol)
    finally:
        c.close()

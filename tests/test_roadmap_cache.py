"""Roadmap reuse across queries of one scene (solve option roadmap_cache).

The reference rebuilds every arm's roadmap for each solve
(pop_planner.cu:812-830), although its own harness runs 50 queries in one
scene (scripts/evaluate_mr.cu:120-140 over scripts/panda_four_problems.json,
whose 50 problems share their obstacles).  With roadmap_cache a context keeps
each arm's roadmap and a later query in the same scene starts from it, its
start and goals inserted as new nodes by a first round.

What is checked here:
* the insertion round against a CPU restatement: the cached nodes are kept,
  the query's nodes follow them, their radius candidates are the query's own
  nodes within range first and then the nearest other nodes, every
  accepted edge is free and every rejected one collides under the FP64
  oracle, and the union-find partition of the whole roadmap equals the
  components of its adjacency;
* results are deterministic for a given sequence of queries, every path is
  valid, and later queries start from the cached nodes and need, on
  average, no more roadmap rounds than the first;
* queries whose start or goal collides do not leave their roadmap behind;
* the server shares the cache between requests.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from mrpop import _lib, geometry, scenes
from oracle import oracle as O
from oracle import roadmap as RM
from test_gpu_plan import check_solution
from test_roadmap import _dump, _uf_roots

pytestmark = pytest.mark.gpu

FOUR = scenes.load_problem_file("tests/golden/panda_four_problems.json", geometry.default_bases(4))


def _st(**kw):
    d = dict(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=2, optimize_iters=300)
    d.update(kw)
    return _lib.settings(**d)


def _op(**kw):
    d = dict(time_budget_s=60.0, roadmap_cache=1)
    d.update(kw)
    return _lib.solve_options_default(**d)


def _components_from_lists(n, deg, nbr):
    edges = set()
    for u in range(n):
        for v in nbr[u][:deg[u]]:
            edges.add((min(u, int(v)), max(u, int(v))))
    return sorted(edges)


@pytest.mark.parametrize("arm", [0, 1, 2, 3])
def test_insertion_round_matches_the_restatement(arm):
    ctx = _lib.Context(0)  # a cache of its own
    try:
        G, rng, kc, K, cap = 8, 2.0, 4, 512, 8192
        st = _st(granularity=G, range=rng)
        op = _op(block_per_robot=1, neighbour_candidates=kc, neighbour_capacity=K, roadmap_capacity=cap)
        first, second = FOUR[0], FOUR[1]
        print(f"[cache] solve {first.name} with the cache", flush=True)
        sol = _lib.solve(ctx, first, st, op)
        assert sol.solved and sol.info["roadmap_cached"] == [0] * 5
        print(f"[cache] {second.name} arm {arm}: insertion round", flush=True)
        q, deg, nbr, uf, connected = _dump(ctx, second, st, op, 1, arm, cap, K)
        n = len(q)
        n_old = n - 1 - len(second.goals)
        assert n_old > 1 + len(first.goals), "no cached nodes"
        assert deg.max() <= K, "a list was truncated: the adjacency would be incomplete"
        # the cached roadmap starts with the first query's start and goal
        fixed = np.vstack([first.start[None], first.goals])[:, 7 * arm:7 * arm + 7]
        assert np.array_equal(q[:1 + len(first.goals)], fixed)
        # the second query's nodes follow the cached ones
        mine = np.vstack([second.start[None], second.goals])[:, 7 * arm:7 * arm + 7].astype(np.float32)
        assert np.array_equal(q[n_old:], mine)
        edges = _components_from_lists(n, deg, nbr)
        # every cached edge was checked when it was made: re-validate them
        old_edges = [(u, v) for u, v in edges if v < n_old]
        ev = O.check_edges(second.scene, O.MODE_SINGLE, second.bases, G, q[[u for u, _ in old_edges]],
                           q[[v for _, v in old_edges]], arm=arm)
        assert not (ev["lo"] & ev["hi"]).any(), "a cached roadmap edge collides"
        # radius candidates of the inserted nodes: the query's nodes within
        # range first, then the nearest others within range (squared float32
        # distance, then id), kc in all
        for k in range(1 + len(second.goals)):
            v = n_old + k
            d = RM.sqdist32(q[:v], q[v])
            r2 = np.float32(np.float32(rng) * np.float32(rng))
            own = [u for u in range(n_old, v) if 0 < d[u] < r2]
            others = sorted((u for u in range(n_old) if 0 < d[u] < r2), key=lambda u: (d[u], u))
            cand = (own + others)[:kc]
            linked = set(int(x) for x in nbr[v][:deg[v]])
            assert linked <= set(cand) | set(range(v + 1, n)), f"node {v}: linked a non-candidate"
            if cand:
                e = O.check_edges(second.scene, O.MODE_SINGLE, second.bases, G, q[cand],
                                  np.repeat(q[v][None], len(cand), 0), arm=arm)
                for u, lo, hi in zip(cand, e["lo"], e["hi"]):
                    got = u in linked
                    assert (not got) or (not lo), f"edge {u}->{v}: colliding edge linked"
                    assert got or hi, f"edge {u}->{v}: free edge not linked"
        # union-find partition of the whole roadmap against its adjacency;
        # start-goal connectivity between the inserted nodes
        root = _uf_roots(uf)
        e = np.array(edges if edges else [(0, 0)])
        g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
        _, lab = connected_components(g, directed=False)
        smallest = {}
        for i, l in enumerate(lab):
            smallest.setdefault(int(l), i)
        assert np.array_equal(root, np.array([smallest[int(l)] for l in lab])), "union-find partition differs"
        assert connected == bool((lab[n_old + 1:n] == lab[n_old]).any()), "rm_connected differs"
        print(f"[cache] arm {arm}: {n_old} cached nodes, {len(edges)} edges, connected {connected}", flush=True)
    finally:
        ctx.close()


def _run_sequence(problems, st, op):
    ctx = _lib.Context(0)
    try:
        return [_lib.solve(ctx, p, st, op) for p in problems]
    finally:
        ctx.close()


def test_cached_results_are_deterministic_and_valid():
    probs = FOUR[:4]
    st, op = _st(), _op()
    print("[cache] sequence 1", flush=True)
    a = _run_sequence(probs, st, op)
    print("[cache] sequence 2", flush=True)
    b = _run_sequence(probs, st, op)
    for p, x, y in zip(probs, a, b):
        check_solution(p, x, granularity=64, rng=2.0)
        assert x.cost == y.cost and np.array_equal(x.path, y.path), f"{p.name}: results differ between runs"
    assert a[0].info["roadmap_cached"][:4] == [0] * 4
    for x in a[1:]:
        assert all(c > 0 for c in x.info["roadmap_cached"][:4]), x.info["roadmap_cached"]
    # later queries start from grown roadmaps: on average they need no more rounds
    assert np.mean([x.info["roadmap_rounds"] for x in a[1:]]) <= a[0].info["roadmap_rounds"]
    print(f"[cache] roadmap rounds {[x.info['roadmap_rounds'] for x in a]}, "
          f"cached nodes {[x.info['roadmap_cached'][:4] for x in a]}", flush=True)
    # without the cache every query builds its own roadmaps
    fresh = _run_sequence(probs, st, _op(roadmap_cache=0))
    for p, x in zip(probs, fresh):
        check_solution(p, x, granularity=64, rng=2.0)
        assert x.info["roadmap_cached"] == [0] * 5


def test_colliding_query_leaves_no_roadmap():
    """An arm whose inserted start collides does not give its roadmap to the
    cache; the other arms do."""
    p = FOUR[0]
    rng = np.random.default_rng(3)
    start = p.start.copy()
    while True:  # a configuration of arm 0 that collides on its own (environment or self)
        q0 = rng.uniform(np.array(geometry.table("kLow")),
                         np.array(geometry.table("kLow")) + np.array(geometry.table("kScale"))).astype(np.float32)
        if O.config_checker(p.scene, 1, p.bases[:1], q0[None])[0]:
            break
    start[:7] = q0
    bad = scenes.Problem(p.name + "_colliding", p.n_arms, p.bases, p.scene, start, p.goals, p.seed)
    ctx = _lib.Context(0)
    try:
        st, op = _st(rrtc_iter=1), _op(time_budget_s=3.0)
        _lib.solve(ctx, bad, st, op)  # solved or not: its arm-0 roadmap must not be kept
        nxt = _lib.solve(ctx, FOUR[1], st, op)
        cached = nxt.info["roadmap_cached"]
        assert cached[0] == 0 and all(c > 0 for c in cached[1:4]), cached
        check_solution(FOUR[1], nxt, granularity=64, rng=2.0)
    finally:
        ctx.close()


def test_server_shares_the_cache():
    ctx = _lib.Context(0)
    try:
        srv = _lib.Server(ctx)
        st, op = _st(), _op()
        t0 = srv.submit(FOUR[0], st, op)
        while srv.poll(t0)[0] != 2:
            srv.step()
        tickets = [srv.submit(p, st, op) for p in FOUR[1:5]]
        while any(srv.poll(t)[0] != 2 for t in tickets):
            srv.step()
        for p, t in zip(FOUR[1:5], tickets):
            _, sol = srv.poll(t, snapshot=True)
            check_solution(p, sol, granularity=64, rng=2.0)
            assert all(c > 0 for c in sol.info["roadmap_cached"][:4]), sol.info["roadmap_cached"]
        srv.close()
    finally:
        ctx.close()

"""Roadmap rounds (P1/P2 of the hot path): the device roadmap after a few
rounds against the CPU restatement in oracle/roadmap.py.

CPU tests pin the generator: Philox4x32-10 known-answer vectors (Random123's
published kat_vectors) for the Python restatement and for the library's
host copy of the device function.  The GPU test rebuilds, round by round,
what the device must contain: the samples (bit-exact), which of them became
nodes (the oracle's single-arm verdicts; near-contact samples may go either
way), the radius candidates of each new node (first Kc of smaller id, float32
distances as on the device), the edge verdicts (oracle, at the granularity),
and every neighbour list in ascending id order truncated to K, with degrees
counting all edges."""
import ctypes

import numpy as np
import pytest

from mrpop import _lib, geometry, scenes
from oracle import oracle as O
from oracle import roadmap as RM

# Random123 kat_vectors, philox4x32 with 10 rounds
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert RM.philox4x32_10(ctr, key) == expect
    lib = _lib.lib()
    U4 = ctypes.c_uint32 * 4
    out = U4()
    lib.mrpop_diag_philox.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.mrpop_diag_philox(ctypes.byref(U4(*ctr)), ctypes.byref((ctypes.c_uint32 * 2)(*key)), ctypes.byref(out))
    assert tuple(out) == expect


def test_u01_range_and_resolution():
    assert RM.u01(0) == np.float32(2.0 ** -24)
    assert RM.u01(0xFFFFFFFF) == np.float32(1.0)


def test_radius_candidates_take_the_first_in_id_order():
    nodes = np.zeros((6, 7), np.float32)
    nodes[:, 0] = [0.1, 5.0, 0.5, 0.9, 0.2, 0.0]
    assert RM.radius_candidates(nodes, 5, 1.0, 8) == [0, 2, 3, 4]
    assert RM.radius_candidates(nodes, 5, 1.0, 2) == [0, 2]
    assert RM.radius_candidates(nodes, 5, 0.9, 8) == [0, 2, 4]  # strictly closer than the range
    nodes[0, 0] = 0.0  # a node repeating node 5's configuration is no neighbour
    assert RM.radius_candidates(nodes, 5, 1.0, 8) == [2, 3, 4]


def test_start_component_matches_connected_components():
    """The reference's label propagation (oracle.roadmap.start_component)
    against scipy's connected components on random graphs, plus known
    answers: a chain reaches its end, a goal on another component does not,
    an isolated start reaches nothing."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    rng = np.random.default_rng(5)
    for trial in range(60):
        n = int(rng.integers(1, 80))
        m = int(rng.integers(0, 2 * n))
        edges = [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(m)]
        edges = [(u, v) for u, v in edges if u != v]
        label0 = RM.start_component(n, edges)
        if edges:
            e = np.array(edges)
            g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
            _, lab = connected_components(g, directed=False)
        else:
            lab = np.arange(n)
        assert np.array_equal(label0, lab == lab[0]), f"trial {trial}"
    chain = [(0, 5), (5, 3), (3, 1)]
    assert RM.start_component(6, chain).tolist() == [True, True, False, True, False, True]
    assert RM.start_goal_connected(RM.start_component(6, chain), 1)
    assert not RM.start_goal_connected(RM.start_component(6, [(0, 5), (1, 2)]), 1)
    assert RM.start_goal_connected(RM.start_component(6, [(0, 5), (5, 2)]), 2)  # the second goal
    assert RM.start_component(4, [(1, 2), (2, 3)]).tolist() == [True, False, False, False]


def _dump(ctx, p, st, op, rounds, arm, cap, K, with_stretch=False):
    lib = _lib.lib()
    lib.mrpop_diag_roadmap.argtypes = ([ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6 +
                                       [ctypes.c_int, ctypes.c_void_p])
    from mrpop._lib import _ProblemRef
    ref = _ProblemRef(p)
    q = np.zeros((cap, 8), np.float32)
    deg = np.zeros(cap, np.int32)
    nbr = np.zeros((cap, K), np.int32)
    uf = np.zeros(cap, np.int32)
    conn = ctypes.c_int()
    n = ctypes.c_int()
    stretch = ctypes.c_int()
    ctx.check(lib.mrpop_diag_roadmap(ctx.handle, ctypes.byref(ref.c), ctypes.byref(st), ctypes.byref(op), rounds, arm,
                                     q.ctypes.data, deg.ctypes.data, nbr.ctypes.data, uf.ctypes.data,
                                     ctypes.byref(conn), ctypes.byref(n), cap, ctypes.byref(stretch)))
    m = n.value
    out = q[:m, :7].copy(), deg[:m].copy(), nbr[:m].copy(), uf[:m].copy(), bool(conn.value)
    return out + (stretch.value,) if with_stretch else out


def _uf_roots(parent):
    """Root of every node of a union-find parent array (host walk)."""
    root = np.empty(len(parent), np.int64)
    for i in range(len(parent)):
        x, steps = i, 0
        while parent[x] != x:
            x = parent[x]
            steps += 1
            assert steps <= len(parent), "union-find parent links form a cycle"
        root[i] = x
    return root


def _check_components(uf, connected, n_nodes, accepted, n_goals):
    """P2: the device's union-find partition and start-goal flag against the
    oracle's label propagation (start component) and scipy's components
    (every component, rooted at its smallest node)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    root = _uf_roots(uf)
    label0 = RM.start_component(n_nodes, accepted)
    assert np.array_equal(root == root[0], label0), "start component differs from the label propagation"
    assert connected == RM.start_goal_connected(label0, n_goals), "rm_connected differs"
    e = np.array(accepted if accepted else [(0, 0)])
    g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n_nodes, n_nodes))
    _, lab = connected_components(g, directed=False)
    smallest = {}
    for i, l in enumerate(lab):
        smallest.setdefault(int(l), i)
    assert np.array_equal(root, np.array([smallest[int(l)] for l in lab])), "union-find partition differs"
    return int(label0.sum()), len(smallest)


def test_biased_sample_indices_follow_the_block_counter():
    """BIASED_ITER bookkeeping (roadmap_interior.cuh:24, 114-117): a block's
    counter is incremented before the test, so with 230 samples per round
    and BIASED_ITER 10000 the first steered sample is iteration 110 of the
    44th round (43 * 230 + 110 = 10000), the next iteration 220 of round 87
    (86 * 230 + 220 = 20000); every block of the round steers the same
    iteration."""
    M = 230 * 8
    for r in range(43):
        assert RM.biased_indices(M, r, 10000) == []
    assert RM.biased_indices(M, 43, 10000) == [109 + 230 * b for b in range(8)]
    assert RM.biased_indices(M, 86, 10000) == [219 + 230 * b for b in range(8)]
    assert RM.biased_indices(M, 86, 0) == []
    assert RM.biased_indices(230, 0, 1) == list(range(230))


def test_steered_sample_geometry():
    """The steered sample against float64 geometry: it lies on the segment
    from its nearest node (brute force over the nodes) towards the sample,
    at distance min(|q - n|, range) from the node, within float32 rounding;
    a sample within range of its nearest node is unchanged."""
    rng = np.random.default_rng(11)
    for trial in range(200):
        nodes = rng.uniform(-3, 3, (int(rng.integers(1, 40)), 7)).astype(np.float32)
        q = rng.uniform(-3, 3, 7).astype(np.float32)
        r = float(rng.uniform(0.2, 3.0))
        d = np.linalg.norm(nodes.astype(np.float64) - q, axis=1)
        nn = nodes[int(np.argmin(d))].astype(np.float64)
        out = RM.steer(q, nodes, r).astype(np.float64)
        v = q - nn
        along = out - nn
        want = min(np.linalg.norm(v), np.float32(r))
        assert abs(np.linalg.norm(along) - want) < 1e-5 * (1 + want), trial
        if np.linalg.norm(v) > 0:
            cos = along @ v / (np.linalg.norm(along) * np.linalg.norm(v))
            assert cos > 1 - 1e-6, trial
        if d.min() <= np.float32(r) * (1 - 1e-6):
            assert np.allclose(out, q, atol=1e-6), trial


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed,arm,n_goals,biased", [
    ("plain2", 4, 0, 1, 10000), ("bin4", 1, 2, 1, 100), ("table3", 3, 0, 1, 10000), ("bin4", 9, 3, 1, 10000),
    ("shelf5", 2, 4, 1, 57), ("shelf5", 6, 1, 1, 10000), ("table3", 4, 1, 3, 31), ("shelf5", 8, 2, 4, 10000)])
def test_roadmap_rounds_match_the_restatement(ctx, kind, seed, arm, n_goals, biased):
    """biased: the steering period (BIASED_ITER); the reference's 10000 is
    not reached in three rounds, so cases with short periods steer samples
    in every round of the first stretch."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx), n_goals=n_goals)
    G, rng, kc, K, bpr, rounds = 8, 2.0, 4, 6, 1, 3
    M = 230 * bpr
    st = _lib.settings(granularity=G, range=rng)
    op = _lib.solve_options_default(block_per_robot=bpr, neighbour_candidates=kc, neighbour_capacity=K,
                                    roadmap_capacity=4096, biased_iter=biased)
    print(f"[roadmap] {kind} seed {seed} arm {arm} goals {n_goals} biased_iter {biased}", flush=True)
    q, deg, nbr, uf, connected, stretch = _dump(ctx, p, st, op, rounds, arm, 4096, K, with_stretch=True)
    assert 1 <= stretch <= rounds  # round 0 always belongs to the first stretch
    scale, low = geometry.table("kScale"), geometry.table("kLow")
    # nodes 0 .. n_goals: start and goal slices of the arm
    n_fixed = 1 + len(p.goals)
    fixed = np.vstack([p.start[None], p.goals])[:, 7 * arm:7 * arm + 7].astype(np.float32)
    assert np.array_equal(q[:n_fixed], fixed)
    nodes = list(fixed)
    accepted = []
    exempt = 0
    moved = 0
    for r in range(rounds):
        samples = RM.roadmap_samples(p.seed, r, arm, M, scale, low)
        if r < stretch:  # rounds of the first stretch: biased samples move towards the roadmap
            arr0 = np.array(nodes, np.float32)
            for i in RM.biased_indices(M, r, biased):
                steered = RM.steer(samples[i], arr0, rng)
                moved += int(not np.array_equal(steered, samples[i]))
                samples[i] = steered
        v = O.check_edges(p.scene, O.MODE_SINGLE, p.bases, 1, samples, samples, arm=arm)
        first_new = len(nodes)
        for i in range(M):
            taken = len(nodes) < len(q) and np.array_equal(q[len(nodes)], samples[i])
            if taken:
                assert not v["lo"][i], f"round {r} sample {i}: a colliding sample became a node"
                nodes.append(samples[i])
            else:
                assert v["hi"][i], f"round {r} sample {i}: a free sample was dropped"
            exempt += int(v["lo"][i] != v["hi"][i])
        arr = np.array(nodes, np.float32)
        frm, to, pairs = [], [], []
        for vtx in range(first_new, len(nodes)):
            for u in RM.radius_candidates(arr, vtx, rng, kc):
                frm.append(arr[u]); to.append(arr[vtx]); pairs.append((u, vtx))
        if pairs:
            ev = O.check_edges(p.scene, O.MODE_SINGLE, p.bases, G, np.array(frm), np.array(to), arm=arm)
            for (u, vtx), lo, hi in zip(pairs, ev["lo"], ev["hi"]):
                # the device's decision: u is in v's list (older neighbours come first, at most kc <= K)
                got = u in nbr[vtx][:min(deg[vtx], K)]
                assert (not got) or (not lo), f"edge {u}->{vtx}: colliding edge linked"
                assert got or hi, f"edge {u}->{vtx}: free edge not linked"
                if got:
                    accepted.append((u, vtx))
    assert len(nodes) == len(q), "the device has nodes the restatement does not"
    lists, degs = RM.expected_lists(len(nodes), accepted, K)
    assert np.array_equal(deg, np.array(degs, np.int32))
    for i, l in enumerate(lists):
        assert list(nbr[i][:len(l)]) == l, f"node {i}: neighbour list differs"
    assert len(accepted) > 100 and max(degs) > K  # the test exercises truncation
    n_start, n_comp = _check_components(uf, connected, len(nodes), accepted, len(p.goals))
    if biased < 10000:
        assert moved > 0, "no sample was steered"
    print(f"[roadmap] {len(nodes)} nodes, {len(accepted)} edges, {exempt} near-contact samples, "
          f"{n_comp} components, {n_start} nodes with the start, connected {connected}, "
          f"{stretch} stretch rounds, {moved} samples steered", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", [("bin4", 3), ("shelf5", 5), ("plain2", 7)])
def test_union_find_matches_the_adjacency(ctx, kind, seed):
    """P2 at planner scale: roadmaps of 1840 samples per round and arm, grown
    for four rounds with neighbour lists long enough to hold every edge; per
    arm the union-find partition, the start component and rm_connected
    against the label propagation and the components of the dumped
    adjacency."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    K, cap = 512, 20000
    st = _lib.settings(granularity=32, range=2.0)
    op = _lib.solve_options_default(block_per_robot=8, neighbour_candidates=64, neighbour_capacity=K,
                                    roadmap_capacity=cap)
    seen_connected = []
    for arm in range(p.n_arms):
        print(f"[components] {kind} seed {seed} arm {arm}", flush=True)
        q, deg, nbr, uf, connected = _dump(ctx, p, st, op, 4, arm, cap, K)
        n = len(q)
        assert deg.max() <= K, "a neighbour list was truncated: the adjacency would be incomplete"
        edges = set()
        for u in range(n):
            for v in nbr[u][:deg[u]]:
                v = int(v)
                assert 0 <= v < n and v != u
                edges.add((min(u, v), max(u, v)))
        for u, v in edges:  # symmetric lists
            assert u in nbr[v][:deg[v]] and v in nbr[u][:deg[u]]
        assert sum(deg) == 2 * len(edges)
        n_start, n_comp = _check_components(uf, connected, n, sorted(edges), len(p.goals))
        seen_connected.append(connected)
        print(f"[components] arm {arm}: {n} nodes, {len(edges)} edges, {n_comp} components, "
              f"{n_start} with the start, connected {connected}", flush=True)
    assert any(seen_connected)

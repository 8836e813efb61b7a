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
    nodes[:, 0] = [0.0, 5.0, 0.5, 0.9, 0.2, 0.0]
    assert RM.radius_candidates(nodes, 5, 1.0, 8) == [0, 2, 3, 4]
    assert RM.radius_candidates(nodes, 5, 1.0, 2) == [0, 2]
    assert RM.radius_candidates(nodes, 5, 0.9, 8) == [0, 2, 4]  # strictly closer than the range


def _dump(ctx, p, st, op, rounds, arm, cap, K):
    lib = _lib.lib()
    lib.mrpop_diag_roadmap.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    from mrpop._lib import _ProblemRef
    ref = _ProblemRef(p)
    q = np.zeros((cap, 8), np.float32)
    deg = np.zeros(cap, np.int32)
    nbr = np.zeros((cap, K), np.int32)
    n = ctypes.c_int()
    ctx.check(lib.mrpop_diag_roadmap(ctx.handle, ctypes.byref(ref.c), ctypes.byref(st), ctypes.byref(op), rounds, arm,
                                     q.ctypes.data, deg.ctypes.data, nbr.ctypes.data, ctypes.byref(n), cap))
    m = n.value
    return q[:m, :7].copy(), deg[:m].copy(), nbr[:m].copy()


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed,arm,n_goals", [("plain2", 4, 0, 1), ("bin4", 1, 2, 1), ("table3", 3, 0, 1),
                                                   ("bin4", 9, 3, 1), ("shelf5", 2, 4, 1), ("shelf5", 6, 1, 1),
                                                   ("table3", 4, 1, 3), ("shelf5", 8, 2, 4)])
def test_roadmap_rounds_match_the_restatement(ctx, kind, seed, arm, n_goals):
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx), n_goals=n_goals)
    G, rng, kc, K, bpr, rounds = 8, 2.0, 4, 6, 1, 3
    M = 230 * bpr
    st = _lib.settings(granularity=G, range=rng)
    op = _lib.solve_options_default(block_per_robot=bpr, neighbour_candidates=kc, neighbour_capacity=K,
                                    roadmap_capacity=4096)
    print(f"[roadmap] {kind} seed {seed} arm {arm} goals {n_goals}", flush=True)
    q, deg, nbr = _dump(ctx, p, st, op, rounds, arm, 4096, K)
    scale, low = geometry.table("kScale"), geometry.table("kLow")
    # nodes 0 .. n_goals: start and goal slices of the arm
    n_fixed = 1 + len(p.goals)
    fixed = np.vstack([p.start[None], p.goals])[:, 7 * arm:7 * arm + 7].astype(np.float32)
    assert np.array_equal(q[:n_fixed], fixed)
    nodes = list(fixed)
    accepted = []
    exempt = 0
    for r in range(rounds):
        samples = RM.roadmap_samples(p.seed, r, arm, M, scale, low)
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
    print(f"[roadmap] {len(nodes)} nodes, {len(accepted)} edges, {exempt} near-contact samples", flush=True)

"""CPU restatement of the planner's roadmap rounds.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

A roadmap round (reference: roadmap_interior.cuh:94-339) for each arm:
  * draws M configurations uniformly in the joint limits (roadmap_interior.cuh:
    105-112); this build draws them from Philox4x32-10 counters keyed by the
    problem seed (the reference's cuRAND streams cannot be reproduced);
  * keeps the collision-free ones (the single-arm checker with the arm's base,
    fkcc_single_buffer_mr, utils.cuh:536-624) as new nodes in sample order;
  * for each new node takes the nodes of smaller id within `range`
    (roadmap_interior.cuh:183-197; here the first `Kc` of them in id order)
    and checks each edge neighbour -> new node at `granularity` samples
    (roadmap_interior.cuh:199-258);
  * links the collision-free edges in both directions;
  * during the solve's first stretch of rounds, steers every BIASED_ITER-th
    sample of a block towards its nearest node (roadmap_interior.cuh:117-165).
Every neighbour list is therefore the node's accepted neighbours in
ascending id order, of which the first `K` are kept; the degree counts all.
Distances are evaluated in float32 exactly as the device does (one fused
multiply-add per coordinate), so the candidate sets are identical.
"""
from __future__ import annotations

import numpy as np

MASK32 = 0xFFFFFFFF
STREAM_ROADMAP = 1
MAX_ARMS = 5


def philox4x32_10(counter, key):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) on 4 x uint32."""
    c0, c1, c2, c3 = (int(v) & MASK32 for v in counter)
    k0, k1 = (int(v) & MASK32 for v in key)
    for _ in range(10):
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        hi0, lo0 = (p0 >> 32) & MASK32, p0 & MASK32
        hi1, lo1 = (p1 >> 32) & MASK32, p1 & MASK32
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + 0x9E3779B9) & MASK32
        k1 = (k1 + 0xBB67AE85) & MASK32
    return c0, c1, c2, c3


def u01(bits: int) -> np.float32:
    """(0, 1] with 24 random bits: ((bits >> 8) + 1) * 2^-24 (exact in float32)."""
    return np.float32(((bits >> 8) + 1) * (1.0 / 16777216.0))


def fma32(a, b, c):
    """float32 fused multiply-add: the product of two float32 and the sum with
    a float32 are exact in float64 for the operands used here, then one
    rounding to float32."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def roadmap_samples(seed: int, rm_round: int, arm: int, M: int, scale, low) -> np.ndarray:
    """The M samples of one roadmap round of one arm, [M, 7] float32."""
    key = (seed & MASK32, (seed >> 32) & MASK32)
    out = np.zeros((M, 7), np.float32)
    a = rm_round * MAX_ARMS + arm
    for i in range(M):
        r0 = philox4x32_10((a, i, 0, STREAM_ROADMAP), key)
        r1 = philox4x32_10((a, i, 1, STREAM_ROADMAP), key)
        bits = (r0[0], r0[1], r0[2], r0[3], r1[0], r1[1], r1[2])
        for j in range(7):
            out[i, j] = fma32(u01(bits[j]), scale[j], low[j])
    return out


ROADMAP_ITER = 230   # samples of one reference block per round (ROADMAP_BUILD_ITER, roadmap_interior.cuh:20)


def biased_indices(M: int, stretch_round: int, biased_iter: int) -> list:
    """Samples of a round that the reference's BIASED_ITER branch steers
    (roadmap_interior.cuh:114-117): sample i is iteration i mod 230 of block
    i // 230, whose counter has reached 230 * stretch_round + (i mod 230) + 1
    (stretch_round = earlier rounds of the solve's first roadmap stretch in
    which the arm sampled)."""
    if biased_iter <= 0:
        return []
    return [i for i in range(M)
            if (stretch_round * ROADMAP_ITER + i % ROADMAP_ITER + 1) % biased_iter == 0]


def steer(sample: np.ndarray, nodes: np.ndarray, rng: float) -> np.ndarray:
    """The biased sample (roadmap_interior.cuh:118-164): its nearest node n
    (squared float32 distance, lowest id among equals) and
    q <- n + (q - n) * min(1, range / |q - n|), each coordinate one fused
    multiply-add; the square root and the division are correctly rounded
    float32 operations."""
    d = sqdist32(nodes, sample)
    bi = int(np.argmin(d))
    dist = np.sqrt(np.float32(d[bi]))
    with np.errstate(divide="ignore"):
        sc = np.float32(min(np.float32(1.0), np.float32(np.float32(rng) / dist)))
    nn = nodes[bi].astype(np.float32)
    return np.array([fma32(np.float32(sample[j] - nn[j]), sc, nn[j]) for j in range(7)], np.float32)


def sqdist32(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Squared distances of rows p [n, 7] to q [7] accumulated as the device
    does: d = fma(x, x, d) over the coordinates, x = p - q in float32.  The
    products are exact in float64; each sum is rounded to float64 and then to
    float32, which differs from the device's single rounding only when the
    float64 sum falls exactly halfway between two float32 values."""
    x = (p.astype(np.float32) - q.astype(np.float32)).astype(np.float64)
    d = np.zeros(len(p), np.float32)
    for j in range(7):
        d = (x[:, j] * x[:, j] + d.astype(np.float64)).astype(np.float32)
    return d


def radius_candidates(nodes: np.ndarray, v: int, rng: float, kc: int) -> list:
    """The first kc nodes of id < v within rng of node v (squared float32
    distance below rng^2, as the device compares); a node at v's own
    configuration is not a neighbour."""
    if v == 0:
        return []
    d = sqdist32(nodes[:v], nodes[v])
    r2 = np.float32(np.float32(rng) * np.float32(rng))
    return [int(u) for u in np.nonzero((d < r2) & (d > 0))[0][:kc]]


def expected_lists(n_nodes: int, accepted_edges, K: int):
    """Neighbour lists (ascending id, first K kept) and degrees from the set
    of accepted undirected edges."""
    nbrs = [[] for _ in range(n_nodes)]
    for u, v in accepted_edges:
        nbrs[u].append(v)
        nbrs[v].append(u)
    lists = [sorted(x) for x in nbrs]
    return [l[:K] for l in lists], [len(l) for l in lists]


def start_component(n_nodes: int, accepted_edges) -> np.ndarray:
    """Nodes in the start's component, by the reference's label propagation
    (roadmap_interior.cuh:275-303): node 0 (the start) carries label 0; every
    sweep gives label 0 to each neighbour of a label-0 node, until a sweep
    changes nothing.  Returns a bool array, True for label 0."""
    adj = [[] for _ in range(n_nodes)]
    for u, v in accepted_edges:
        adj[u].append(v)
        adj[v].append(u)
    label0 = np.zeros(n_nodes, bool)
    if n_nodes:
        label0[0] = True
    change = True
    while change:
        change = False
        for t in range(n_nodes):
            if label0[t]:
                for nn in adj[t]:
                    if not label0[nn]:
                        label0[nn] = True
                        change = True
    return label0


def start_goal_connected(label0: np.ndarray, n_goals: int) -> bool:
    """Some goal node 1..n_goals carries the start's label
    (roadmap_interior.cuh:305-313)."""
    return bool(label0[1:1 + n_goals].any())

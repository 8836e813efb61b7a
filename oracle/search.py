"""CPU restatement of one dRRT search round of the planner.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

A round (reference: pop_planner.cu:350-736, made round-synchronous: every
expansion of a round sees the trees as they stood at its start) is, for
each expansion b of the B of a round:
  * tree choice (pop_planner.cu:366-376): the first round of a launch splits
    b over the trees (b < B/2: tree 0); later rounds keep b's tree, flipped
    to the other one when |n_t - n_o| / n_t < tree_ratio (balance 2);
  * a uniform sample in the joint box (pop_planner.cu:389-397), drawn from
    Philox4x32-10 counters (round, b, lane / 4, stream 2) keyed by the seed;
  * the nearest node of the tree (pop_planner.cu:472-508): least float32
    squared distance, the lowest index among equals;
  * per arm, the roadmap neighbour of that node's roadmap node, other than
    itself, within `range`, of least angle between (sample - node) and
    (neighbour - node) (utils.cuh:688-759, 62-95); no neighbour: no
    expansion;
  * an existing node within squared distance 1e-5 of the new composite
    configuration: no expansion (pop_planner.cu:526-549, first pass);
  * the candidate edge near -> new, checked once per distinct (tree, near,
    roadmap nodes) in the round; free candidates are inserted in expansion
    order (pop_planner.cu:551-622) with cost(near) + |sample - near|;
  * for each inserted node, the nearest node of the other tree and the
    straight connection in n = floor(d / range) + 1 chunks; a connection
    whose chunks are all free adds its end node (cost: one chunk's length,
    pop_planner.cu:624-731), and the lowest such expansion records the path
    goal -> start (path_processing.cu:10-70).

Optimise passes (AO-x pass >= 1, pop_planner.cu:389-469) draw the sample
from the prolate hyperspheroid around start and goal, reject samples whose
distance sum reaches the best cost, bound node costs by c_rand (nearest
node among cheaper nodes, roadmap steps within the remaining bound) and
rewire an existing node when the new cost is lower; phs_sample64 and
cost_bound64 restate the sampler and the bound in float64.

Float32 arithmetic is reproduced exactly where the device's is exact:
fused multiply-adds (fma32, below), correctly rounded division and square
root, single additions.  The device's acosf is not reproduced: angles are
compared in float64, and the test accepts the device's pick when no other
neighbour is smaller by more than ANGLE_TOL.
"""
from __future__ import annotations

import numpy as np

from .roadmap import MASK32, philox4x32_10, u01

STREAM_TREE = 2
STREAM_COST_BOUND = 3
STREAM_PHS_RADIUS = 4
FLT_MAX = np.float32(np.finfo(np.float32).max)
NONE = 0x7FFFFFFF
ANGLE_TOL = 1e-5  # radians


def fma32(a, b, c):
    """float32 fma(a, b, c) with one rounding, elementwise.  The float64
    product is exact; the float64 sum and its exact error (TwoSum) decide
    the float32 rounding, including sums that fall on a float32 midpoint."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    c = np.asarray(c, np.float32).astype(np.float64)
    p = a * b
    s = p + c
    bb = s - p
    err = (p - (s - bb)) + (c - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    toward = np.where(s > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    other = np.nextafter(r, toward)
    mid = (s != r64) & ((r64 + other.astype(np.float64)) * 0.5 == s)
    lo = np.minimum(r, other)
    hi = np.maximum(r, other)
    r = np.where(mid & (err > 0), hi, np.where(mid & (err < 0), lo, r))
    return r.astype(np.float32)


def sqdist(nodes: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Squared distances of rows `nodes` [n, D] to q [D] as the device sums
    them: x = node - q in float32, s = fma(x, x, s) over the coordinates."""
    nodes = np.atleast_2d(np.asarray(nodes, np.float32))
    x = nodes - np.asarray(q, np.float32)[None, :]
    s = np.zeros(len(nodes), np.float32)
    for d in range(nodes.shape[1]):
        s = fma32(x[:, d], x[:, d], s)
    return s


def sqrt32(x):
    return np.sqrt(np.asarray(x, np.float32)).astype(np.float32)


def angle64(v, q, r) -> float:
    """Angle at v between q - v and r - v (utils.cuh:62-95), in float64."""
    a = np.asarray(q, np.float64) - np.asarray(v, np.float64)
    b = np.asarray(r, np.float64) - np.asarray(v, np.float64)
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    if na < 1e-8 or nb < 1e-8:
        return 0.0
    return float(np.arccos(np.clip(a @ b / (na * nb), -1.0, 1.0)))


def tree_sample(seed: int, rng_round: int, b: int, D: int, scale, low) -> np.ndarray:
    key = (seed & MASK32, (seed >> 32) & MASK32)
    q = np.zeros(D, np.float32)
    for lane in range(D):
        r = philox4x32_10((rng_round, b, lane >> 2, STREAM_TREE), key)
        q[lane] = fma32(u01(r[lane & 3]), scale[lane % 7], low[lane % 7])
    return q


class Slot:
    """The state a round reads: trees (tr_q [2][cap][D], tr_parent,
    tr_cost, tr_rid [2][cap][A], tr_n [2]), roadmaps (rm_q [A][cap][8],
    rm_nbr, rm_deg), the previous round's tree choices (ex_tree) and the
    round counters."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def tree_choice(s: Slot, b: int) -> int:
    t = (0 if b < s.B // 2 else 1) if s.round == 0 else int(s.ex_tree[b])
    if s.round > 0 and s.balance == 2:
        nt, no = int(s.tr_n[t]), int(s.tr_n[1 - t])
        ratio = np.float32(abs(nt - no)) / np.float32(nt)
        if ratio < np.float32(s.tree_ratio):
            t = 1 - t
    return t


def phs_sample64(s: Slot, b: int) -> np.ndarray:
    """The optimise-pass sample of expansion b in float64 (pop_planner.cu:
    398-442): the logistic quantile of each uniform times sqrt(pi/8) as an
    approximate normal, normalised, scaled by u^(1/D), mapped by the
    hyperspheroid transform (phs_tf, phs_center), clamped to the joint box.
    The device evaluates logf/powf in float32; compare with a tolerance."""
    key = (s.seed & MASK32, (s.seed >> 32) & MASK32)
    D = s.D
    z = np.zeros(D)
    for lane in range(D):
        r = philox4x32_10((s.rng_round, b, lane >> 2, STREAM_TREE), key)
        u = min(float(u01(r[lane & 3])), 1.0 - 2.0 ** -24)
        z[lane] = np.log(u / (1.0 - u)) * np.sqrt(np.pi / 8.0)
    z /= np.linalg.norm(z)
    rad = float(u01(philox4x32_10((s.rng_round, b, 0, STREAM_PHS_RADIUS), key)[0]))
    v = s.phs_tf.astype(np.float64) @ (z * rad ** (1.0 / D)) + s.phs_center.astype(np.float64)
    sc = np.tile(s.scale.astype(np.float64), D // 7)
    lo = np.tile(s.low.astype(np.float64), D // 7)
    return np.clip((v - lo) / sc, 0.0, 1.0) * sc + lo


def cost_bound64(s: Slot, b: int, t: int, q) -> float:
    """c_rand (pop_planner.cu:459-469) in float64."""
    key = (s.seed & MASK32, (s.seed >> 32) & MASK32)
    q = np.asarray(q, np.float64)
    g = np.linalg.norm(s.tr_q[t, 0].astype(np.float64) - q)
    h = np.linalg.norm(s.tr_q[1 - t, 0].astype(np.float64) - q)
    c_range = max(float(s.best) - (g + h), 0.0)
    return float(u01(philox4x32_10((s.rng_round, b, 0, STREAM_COST_BOUND), key)[0])) * c_range + g


def expand(s: Slot, b: int, q=None, c_rand=None):
    """Expansion b of the round up to the roadmap step.  Returns (tree,
    near, dnear, cands, bound) with cands[a] the admissible neighbours of
    arm a as (angle, id, distance) sorted by angle, or None when an arm has
    none.  First pass: q from the Philox stream.  Optimise passes: q and
    c_rand are the device's (checked against phs_sample64 / cost_bound64),
    the nearest node is taken among nodes cheaper than c_rand and roadmap
    neighbours farther than the remaining cost bound are not admissible."""
    t = tree_choice(s, b)
    if q is None:
        q = tree_sample(s.seed, s.rng_round, b, s.D, s.scale, s.low)
    nt = int(s.tr_n[t])
    d = sqdist(s.tr_q[t, :nt], q)
    if c_rand is not None:
        d = np.where(s.tr_cost[t, :nt] < np.float32(c_rand), d, np.float32(np.inf))
    near = int(np.argmin(d))
    dnear = sqrt32(d[near])
    vn = s.tr_q[t, near]
    bound = np.float32(c_rand) - s.tr_cost[t, near] if c_rand is not None else None
    cands = []
    for a in range(s.A):
        rid = int(s.tr_rid[t, near, a])
        deg = min(int(s.rm_deg[a, rid]), s.K)
        c = []
        for r in s.rm_nbr[a, rid, :deg]:
            r = int(r)
            if r == rid:
                continue
            p = s.rm_q[a, r, :7]
            d2 = sqdist(vn[None, 7 * a:7 * a + 7], p)[0]
            dr = sqrt32(d2)
            if dr > np.float32(s.range) or d2 == 0:  # a node repeating the arm's configuration is no move
                continue
            c.append((angle64(vn[7 * a:7 * a + 7], q[7 * a:7 * a + 7], p), r, dr))
        if not c:
            return t, near, dnear, None, bound
        cands.append(sorted(c))
    return t, near, dnear, cands, bound


def pick_ok(cands_a, r: int) -> bool:
    """The device's pick r for one arm: admissible, and no neighbour's angle
    is smaller by more than ANGLE_TOL (the device compares acosf values)."""
    ang = dict((c[1], c[0]) for c in cands_a)
    return r in ang and ang[r] <= cands_a[0][0] + ANGLE_TOL


def compose(s: Slot, t: int, near: int, dnear, nrid):
    """New configuration, its cost (cost(near) + |sample - near|) and the
    first node of tree t within squared distance 1e-5 of it (None if none)."""
    nq = np.concatenate([s.rm_q[a, nrid[a], :7] for a in range(s.A)]).astype(np.float32)
    cost = np.float32(s.tr_cost[t, near] + dnear)
    nt = int(s.tr_n[t])
    close = np.nonzero(sqdist(s.tr_q[t, :nt], nq) <= np.float32(1e-5))[0]
    return nq, cost, (int(close[0]) if len(close) else None)


def connect_plan(s: Slot, t: int, q: np.ndarray, ce_per_b: int, ub=None):
    """Nearest node of the other tree and the connection chunks
    (pop_planner.cu:624-703), or None when the connection has more chunks
    than an expansion may use.  Optimise passes: only nodes closer than ub
    (best cost - cost of the new node) count; None if there is none."""
    o = 1 - t
    no = int(s.tr_n[o])
    d2 = sqdist(s.tr_q[o, :no], q)
    if ub is not None:
        d2 = np.where(sqrt32(d2) < np.float32(ub), d2, np.float32(np.inf))
        if not np.isfinite(d2).any():
            return None
    on = int(np.argmin(d2))
    d = sqrt32(d2[on])
    n = int(np.float32(d) / np.float32(s.range)) + 1
    if n > ce_per_b:
        return None
    vec = ((s.tr_q[o, on] - q) / np.float32(n)).astype(np.float32)
    frm, to = [], []
    cur = q.astype(np.float32).copy()
    for _ in range(n):
        frm.append(cur.copy())
        cur = (cur + vec).astype(np.float32)
        to.append(cur.copy())
    return dict(near=on, next=n, d=d, frm=np.array(frm), to=np.array(to), end=cur)


def record_path(tr_q, tr_parent, t: int, wnode: int, onode: int, path_cap: int):
    """Goal -> start path through the connection (path_processing.cu:10-70):
    tree 1's chain reversed, then tree 0's chain to its root; cost = float32
    sum of segment lengths, tree 1 first.  The reference writes node 0 of each
    tree as the root (path_processing.cu:31-33, 48-50): the same node for the
    start tree and for a single goal; with several goals the goal reached is
    written here, not goal 0."""
    start = {t: wnode, 1 - t: onode}
    chains = []
    for tr in (0, 1):
        cur, chain = start[tr], [start[tr]]
        while tr_parent[tr, cur] != cur and len(chain) <= path_cap:
            cur = int(tr_parent[tr, cur])
            chain.append(cur)
        chains.append(chain)
    cost = np.float32(0)
    for tr in (1, 0):
        ch = chains[tr]
        for a, b in zip(ch[:-1], ch[1:]):
            cost = np.float32(cost + sqrt32(sqdist(tr_q[tr, a][None], tr_q[tr, b])[0]))
    path = np.concatenate([tr_q[1, chains[1][::-1]], tr_q[0, chains[0]]])
    return path, cost

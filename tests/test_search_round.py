"""dRRT search rounds (P3-P8 of the hot path): the device's rounds against
the CPU restatement in oracle/search.py.

A problem's slot is driven round by round through the library's diagnostic
entry points; before each round its trees, roadmaps and counters are read,
after it every expansion record, candidate edge, connection chunk and the
trees again.  The restatement recomputes the round from the state before it:
tree choice and balance, the Philox uniform sample (bit-exact), the nearest
node (float32 distances, lowest index among equals), the roadmap step per
arm (the device's pick must be admissible and of least angle within
ANGLE_TOL; acosf is not reproduced), the existing-node test, the new cost,
one checked edge per distinct (tree, near, roadmap nodes), insertion in
expansion order, the connection to the other tree in floor(d / range) + 1
chunks, the connection node and the recorded path.  The edge verdicts the
device used are checked against the FP64 oracle (near-contact edges may go
either way) and then drive the restated commit.

CPU tests pin fma32, the exact float32 fused multiply-add the restatement
relies on, and the path record on a hand-built two-goal tree."""
import ctypes
import dataclasses
from fractions import Fraction

import numpy as np
import pytest

from mrpop import _lib, geometry, scenes
from oracle import oracle as O
from oracle import search as SR
from test_gpu_plan import interior_waypoint
import test_phs


def _round_half_even_f32(x: Fraction) -> np.float32:
    """Correct rounding of an exact rational to float32 (normal range)."""
    if x == 0:
        return np.float32(0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    m = x / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return np.float32(sign * float(Fraction(fl) * Fraction(2) ** (e - 23)))


def test_fma32_is_one_rounding():
    rng = np.random.default_rng(0)
    a = rng.standard_normal(4000).astype(np.float32)
    b = rng.standard_normal(4000).astype(np.float32)
    c = (rng.standard_normal(4000) * 10.0 ** rng.integers(-8, 3, 4000)).astype(np.float32)
    got = SR.fma32(a, b, c)
    want = [_round_half_even_f32(Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z)))
            for x, y, z in zip(a, b, c)]
    assert np.array_equal(got, np.array(want, np.float32))
    # x * x = 1 + 2^-11 + 2^-24 lies halfway between two floats; an addend
    # below float64 resolution vanishes from the float64 sum, which then sits
    # on the midpoint, but it decides the one correct rounding
    x = np.float32(1 + 2.0 ** -12)
    for c in (2.0 ** -60, -(2.0 ** -60), 0.0):
        c = np.float32(c)
        want = _round_half_even_f32(Fraction(float(x)) ** 2 + Fraction(float(c)))
        assert SR.fma32(x, x, c) == want, c
    assert SR.fma32(x, x, np.float32(2.0 ** -60)) > SR.fma32(x, x, np.float32(-(2.0 ** -60)))


def test_sqdist_accumulates_with_fma():
    rng = np.random.default_rng(1)
    p = rng.standard_normal((50, 14)).astype(np.float32)
    q = rng.standard_normal(14).astype(np.float32)
    d = SR.sqdist(p, q)
    for i in range(50):
        s = np.float32(0)
        for k in range(14):
            x = np.float32(p[i, k] - q[k])
            s = _round_half_even_f32(Fraction(float(x)) ** 2 + Fraction(float(s)))
        assert d[i] == s


def test_record_path_on_a_two_goal_tree():
    """Known answer on hand-built trees (one dimension, so lengths are
    differences): tree 1 has goal roots 0 and 1 and the chain 3 -> 2 -> 1;
    tree 0 the chain 2 -> 1 -> 0.  The path runs from goal 1 (the root the
    chain reached, not goal 0) through the connection to the start."""
    tr_q = np.zeros((2, 4, 1), np.float32)
    tr_q[1, :, 0] = [10.0, 9.0, 7.0, 4.0]    # goals 10 and 9; nodes 7 and 4
    tr_q[0, :, 0] = [0.0, 1.0, 3.0, 0.0]     # start 0; nodes 1 and 3
    tr_parent = np.array([[0, 0, 1, 0], [0, 1, 1, 2]])
    for t, wnode, onode in ((1, 3, 2), (0, 2, 3)):  # the same connection seen from either tree
        path, cost = SR.record_path(tr_q, tr_parent, t, wnode, onode, 16)
        assert path[:, 0].tolist() == [9.0, 7.0, 4.0, 3.0, 1.0, 0.0]
        # tree 1's segments, then tree 0's; the joining segment 4 -> 3 is not
        # counted (in the planner its two ends are the same configuration)
        assert cost == np.float32(2 + 3 + 2 + 1)


# ---------------------------------------------------------------- GPU ----

# Expansion forms: "wave" = one wave per expansion; "split" = the five-launch
# form from the first round, with tree slices of 64 rows so that even these
# small trees are scanned by several blocks per expansion group.  The default
# switches from the first to the second at 1024 tree nodes.
_FORMS = {"wave": (-1, 256), "split": (0, 64), "default": (1024, 256)}


class _expand_form:
    def __init__(self, ctx, form):
        self.ctx, self.form = ctx, form
        self.lib = _lib.lib()
        self.lib.mrpop_diag_set_expand_split.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]

    def __enter__(self):
        self.ctx.check(self.lib.mrpop_diag_set_expand_split(self.ctx.handle, *_FORMS[self.form]))

    def __exit__(self, *exc):
        self.ctx.check(self.lib.mrpop_diag_set_expand_split(self.ctx.handle, *_FORMS["default"]))

class _DiagSlot:
    def __init__(self, ctx, problem, st, op):
        self.lib = _lib.lib()
        L = self.lib
        L.mrpop_diag_slot_open.argtypes = [ctypes.c_void_p] * 5
        L.mrpop_diag_slot_roadmap.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.mrpop_diag_slot_search.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.mrpop_diag_slot_info.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mrpop_diag_slot_read.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong]
        L.mrpop_diag_slot_read.restype = ctypes.c_longlong
        L.mrpop_diag_slot_close.argtypes = [ctypes.c_void_p]
        L.mrpop_diag_slot_close.restype = None
        self.ctx = ctx
        self.ref = _lib._ProblemRef(problem)
        self.st, self.op = st, op
        self.h = ctypes.c_void_p()
        ctx.check(L.mrpop_diag_slot_open(ctx.handle, ctypes.byref(self.ref.c), ctypes.byref(st), ctypes.byref(op),
                                         ctypes.byref(self.h)))
        L.mrpop_diag_slot_simplify.argtypes = [ctypes.c_void_p, ctypes.c_int]
        info = np.zeros(11, np.int32)
        ctx.check(L.mrpop_diag_slot_info(self.h, info.ctypes.data))
        (self.A, self.D, self.B, self.M, self.rm_cap, self.K, self.tr_cap, self.ce_cap, self.G, self.path_cap,
         self.sp_cap) = (int(v) for v in info)

    def close(self):
        self.lib.mrpop_diag_slot_close(self.h)

    def roadmap(self, rounds):
        self.ctx.check(self.lib.mrpop_diag_slot_roadmap(self.h, rounds))

    def search(self, rounds, keep_round):
        self.ctx.check(self.lib.mrpop_diag_slot_search(self.h, rounds, int(keep_round)))

    def simplify(self, stage):
        self.ctx.check(self.lib.mrpop_diag_slot_simplify(self.h, stage))

    def start_pass(self, k):
        self.lib.mrpop_diag_slot_pass.argtypes = [ctypes.c_void_p, ctypes.c_int]
        self.ctx.check(self.lib.mrpop_diag_slot_pass(self.h, k))

    def read(self, name, dtype, shape):
        n = int(np.prod(shape))
        out = np.zeros(n, dtype)
        got = self.lib.mrpop_diag_slot_read(self.h, name.encode(), out.ctypes.data, out.nbytes)
        assert got >= out.nbytes, (name, got, out.nbytes)
        return out.reshape(shape)

    def state(self):
        A, D, B, tc, rc, K = self.A, self.D, self.B, self.tr_cap, self.rm_cap, self.K
        ctr = self.read("counters", np.int32, (64,))
        return dict(
            counters=ctr, tr_n=ctr[0:2].copy(), status=int(ctr[4]),
            best=ctr[5:6].view(np.float32)[0], path_cost=ctr[6:7].view(np.float32)[0], path_n=int(ctr[7]),
            round=int(ctr[16]), rng_round=int(ctr[17]),
            tr_q=self.read("tr_q", np.float32, (2, tc, D)), tr_parent=self.read("tr_parent", np.int32, (2, tc)),
            tr_cost=self.read("tr_cost", np.float32, (2, tc)), tr_rid=self.read("tr_rid", np.int32, (2, tc, A)),
            ex_tree=self.read("ex_tree", np.int32, (B,)), ex_state=self.read("ex_state", np.int32, (B,)),
            ex_near=self.read("ex_near", np.int32, (B,)), ex_dup=self.read("ex_dup", np.int32, (B,)),
            ex_index=self.read("ex_index", np.int32, (B,)), ex_cost=self.read("ex_cost", np.float32, (B,)),
            ex_q=self.read("ex_q", np.float32, (B, D)), ex_rid=self.read("ex_rid", np.int32, (B, A)),
            ed_b=self.read("ed_b", np.int32, (B,)), ed_coll=self.read("ed_coll", np.uint32, (B,)),
            ed_from=self.read("ed_from", np.float32, (B, D)), ed_to=self.read("ed_to", np.float32, (B, D)),
            cn_near=self.read("cn_near", np.int32, (B,)), cn_next=self.read("cn_next", np.int32, (B,)),
            cn_d=self.read("cn_d", np.float32, (B,)), cn_fail=self.read("cn_fail", np.uint32, (B,)),
            ce_count=int(ctr[3]), ce_b=self.read("ce_b", np.int32, (self.ce_cap,)),
            ce_coll=self.read("ce_coll", np.uint32, (self.ce_cap,)),
            ce_from=self.read("ce_from", np.float32, (self.ce_cap, D)),
            ce_to=self.read("ce_to", np.float32, (self.ce_cap, D)),
            path=self.read("path", np.float32, (self.path_cap, D)),
            ex_sample=self.read("ex_sample", np.float32, (B, D)), ex_crand=self.read("ex_crand", np.float32, (B,)),
            rm_q=self.read("rm_q", np.float32, (A, rc, 8)), rm_nbr=self.read("rm_nbr", np.int32, (A, rc, K)),
            rm_deg=self.read("rm_deg", np.int32, (A, rc)))


def _verdicts_agree(problem, mode, G, frm, to, device):
    """Device edge verdicts against the FP64 oracle; near-contact edges are
    exempt.  Returns the number of exempt edges."""
    if not len(frm):
        return 0
    r = O.check_edges(problem.scene, mode, problem.bases, G, np.array(frm), np.array(to))
    device = np.asarray(device, bool)
    robust = r["lo"] == r["hi"]
    assert np.array_equal(device[robust], r["lo"][robust]), f"mode {mode}: edge verdicts differ from the oracle"
    return int((~robust).sum())


def _check_round(p, ds, before, after, settings, optimize=False, phs=None):
    """Restate one round from `before` and compare with `after`.  First
    pass: everything from the state before the round.  Optimise passes: the
    sample and its cost bound are the device's, checked in float64 against
    the hyperspheroid sampler and the bound formula; the informed rejection,
    the cost-filtered nearest node, the bounded roadmap step, rewires and
    inserts are restated from them."""
    A, D, B = ds.A, ds.D, ds.B
    s = SR.Slot(A=A, D=D, B=B, K=ds.K, seed=p.seed, range=settings.range, balance=settings.balance,
                tree_ratio=settings.tree_ratio, round=before["round"], rng_round=before["rng_round"],
                scale=geometry.table("kScale"), low=geometry.table("kLow"), tr_n=before["tr_n"],
                tr_q=before["tr_q"], tr_cost=before["tr_cost"], tr_rid=before["tr_rid"], ex_tree=before["ex_tree"],
                rm_q=before["rm_q"], rm_nbr=before["rm_nbr"], rm_deg=before["rm_deg"], best=before["best"])
    assert before["status"] == 0
    if not optimize:
        assert before["best"] == SR.FLT_MAX  # first pass, still searching
    else:
        s.phs_tf, s.phs_center = phs
    best = before["best"]
    cand = {}
    rejected = 0
    inside = 0
    for b in range(B):
        t = SR.tree_choice(s, b)
        assert after["ex_tree"][b] == t, f"expansion {b}: tree"
        q = c_rand = None
        if optimize:
            q = after["ex_sample"][b]
            q64 = SR.phs_sample64(s, b)
            assert np.abs(q.astype(np.float64) - q64).max() < 1e-4, f"expansion {b}: PHS sample"
            g = (np.linalg.norm(s.tr_q[0, 0].astype(np.float64) - q) + np.linalg.norm(s.tr_q[1, 0].astype(np.float64) - q))
            # the sample lies in the hyperspheroid with foci start and goal 0 and
            # transverse diameter best (pop_planner.cu:398-442): the transform
            # maps the unit ball into it, and clamping to the joint box, a
            # projection onto a convex set holding both foci, moves no point
            # away from either focus.  When a path to a nearer goal made best
            # shorter than the focal distance, the hyperspheroid is the
            # segment between the foci and the bound cannot hold.
            cmin = float(np.linalg.norm(s.tr_q[1, 0].astype(np.float64) - s.tr_q[0, 0]))
            if float(best) >= cmin:
                assert g <= float(best) * (1 + 1e-5) + 1e-5, f"expansion {b}: PHS sample outside the hyperspheroid"
                inside += 1
            dev_rejected = after["ex_crand"][b] == -1.0
            if abs(g - float(best)) > 1e-5:
                assert dev_rejected == (g >= best), f"expansion {b}: informed rejection"
            # else: within float rounding of the bound (a degenerate hyperspheroid
            # when the best path is the straight segment): the device decides
            if dev_rejected:
                rejected += 1
                continue
            c_rand = after["ex_crand"][b]
            assert abs(float(c_rand) - SR.cost_bound64(s, b, t, q)) < 1e-4, f"expansion {b}: cost bound"
        t, near, dnear, cands, bound = SR.expand(s, b, q, c_rand)
        if cands is None:
            continue
        # least angle among the admissible (within the cost bound left, which
        # each arm's step lowers); among angles within ANGLE_TOL of it the
        # device's pick
        nrid = []
        vn = s.tr_q[t, near]
        for a in range(A):
            adm = [c for c in cands[a] if bound is None or not (c[2] > bound)]
            if not adm:
                break
            r0 = adm[0][1]
            tie = len(adm) > 1 and adm[1][0] <= adm[0][0] + SR.ANGLE_TOL
            dev = int(after["ex_rid"][b][a])
            nrid.append(dev if tie and SR.pick_ok(adm, dev) else r0)
            if bound is not None:
                step = SR.sqrt32(SR.sqdist(s.rm_q[a, nrid[-1], :7][None], vn[7 * a:7 * a + 7])[0])
                bound = np.float32(bound - step)
        if len(nrid) < A:
            continue
        nq, cost, exist = SR.compose(s, t, near, dnear, nrid)
        state = 1
        if exist is not None:
            if not optimize or not cost < s.tr_cost[t, exist]:
                continue
            state = 2
        if optimize and not cost < best:
            continue
        cand[b] = dict(t=t, near=near, nrid=tuple(nrid), nq=nq, cost=cost, state=state, exist=exist)
    # candidates: one edge per distinct (tree, near, roadmap nodes)
    groups = {}
    for b, c in cand.items():
        groups.setdefault((c["t"], c["near"], c["nrid"]), []).append(b)
    n_edges = len(groups)
    emitted = after["ed_b"][:n_edges].tolist()
    assert sorted(set(emitted)) == sorted(emitted) and len(emitted) == n_edges
    for b in range(B):
        if b not in cand:
            assert after["ex_dup"][b] == -1 and after["ex_state"][b] == 0 and b not in emitted, \
                f"expansion {b}: the device made a candidate of it"
            assert after["ex_index"][b] == -1
    verdict = {}
    frm, to, dev_v = [], [], []
    for key, members in groups.items():
        owners = [b for b in members if b in emitted]
        assert len(owners) == 1, f"group {key}: {len(owners)} edges"
        w = owners[0]
        for b in members:
            assert after["ex_dup"][b] == (-1 if b == w else w), f"expansion {b}: duplicate link"
            c = cand[b]
            assert after["ex_near"][b] == c["near"] and tuple(after["ex_rid"][b]) == c["nrid"]
            assert np.array_equal(after["ex_q"][b], c["nq"]) and after["ex_cost"][b] == c["cost"], f"expansion {b}"
        e = emitted.index(w)
        c = cand[w]
        assert np.array_equal(after["ed_from"][e], before["tr_q"][c["t"], c["near"]])
        assert np.array_equal(after["ed_to"][e], c["nq"])
        frm.append(after["ed_from"][e]); to.append(after["ed_to"][e]); dev_v.append(after["ed_coll"][e] != 0)
        for b in members:
            verdict[b] = after["ed_coll"][e] != 0
    exempt = _verdicts_agree(p, O.MODE_DRRT, ds.G, frm, to, dev_v)
    # commit: rewires (lowest cost, then lowest expansion, per existing node),
    # then free inserts in expansion order per tree
    tr_q, tr_parent = before["tr_q"].copy(), before["tr_parent"].copy()
    tr_cost, tr_rid, tr_n = before["tr_cost"].copy(), before["tr_rid"].copy(), before["tr_n"].copy()
    rewires = {}
    for b in sorted(cand):
        c = cand[b]
        if c["state"] == 2 and not verdict[b]:
            k = (c["t"], c["exist"])
            if k not in rewires or (c["cost"], b) < (cand[rewires[k]]["cost"], rewires[k]):
                rewires[k] = b
    for (t, x), b in rewires.items():
        tr_parent[t, x], tr_cost[t, x] = cand[b]["near"], cand[b]["cost"]
    index = {}
    for b in sorted(cand):
        if verdict[b]:
            assert after["ex_state"][b] == 0 and after["ex_index"][b] == -1, f"expansion {b}: colliding edge committed"
            continue
        c = cand[b]
        if c["state"] == 2:
            assert after["ex_state"][b] == 2 and after["ex_index"][b] == -1, f"expansion {b}: rewire"
            continue
        t, idx = c["t"], int(tr_n[c["t"]])
        assert idx < ds.tr_cap
        tr_n[t] += 1
        tr_q[t, idx], tr_parent[t, idx], tr_cost[t, idx], tr_rid[t, idx] = c["nq"], c["near"], c["cost"], c["nrid"]
        index[b] = idx
        assert after["ex_state"][b] == 1 and after["ex_index"][b] == idx, f"expansion {b}: insert"
    # connection of each inserted node to the other tree (trees after the commit)
    s2 = SR.Slot(**dict(s.__dict__, tr_n=tr_n, tr_q=tr_q))
    plans = {}
    cfrm, cto, cdev = [], [], []
    ce_b = after["ce_b"][:after["ce_count"]].tolist()
    for b in range(B):
        if b not in index:
            assert after["cn_near"][b] == -1
            continue
        t = cand[b]["t"]
        ub = np.float32(best - tr_cost[t, index[b]]) if optimize else None
        plan = SR.connect_plan(s2, t, tr_q[t, index[b]], ds.ce_cap // B, ub)
        if plan is None:
            assert after["cn_near"][b] == -1
            continue
        assert after["cn_near"][b] == plan["near"] and after["cn_next"][b] == plan["next"]
        assert after["cn_d"][b] == plan["d"]
        base = ce_b.index(b)
        assert ce_b[base:base + plan["next"]] == [b] * plan["next"]
        assert np.array_equal(after["ce_from"][base:base + plan["next"]], plan["frm"])
        assert np.array_equal(after["ce_to"][base:base + plan["next"]], plan["to"])
        chunks = after["ce_coll"][base:base + plan["next"]] != 0
        cfrm += list(plan["frm"]); cto += list(plan["to"]); cdev += list(chunks)
        plan["failed"] = bool(chunks.any())
        assert bool(after["cn_fail"][b]) == plan["failed"]
        plans[b] = plan
    exempt += _verdicts_agree(p, O.MODE_FULL, ds.G, cfrm, cto, cdev)
    assert len(ce_b) == sum(pl["next"] for pl in plans.values())
    # connection nodes, in expansion order per tree; the lowest records the path
    n_commit = tr_n.copy()
    winner = None
    for b in sorted(plans):
        pl = plans[b]
        if pl["failed"]:
            continue
        t, o = cand[b]["t"], 1 - cand[b]["t"]
        idx = int(tr_n[t])
        tr_n[t] += 1
        tr_q[t, idx] = pl["end"]
        tr_rid[t, idx] = tr_rid[o, pl["near"]]
        tr_parent[t, idx] = index[b]
        tr_cost[t, idx] = np.float32(tr_cost[t, index[b]] + np.float32(pl["d"] / np.float32(pl["next"])))
        if winner is None:
            winner = (b, t, idx, pl["near"])
    for tr in (0, 1):
        n = int(tr_n[tr])
        assert after["tr_n"][tr] == n, f"tree {tr}: {after['tr_n'][tr]} nodes, restated {n}"
        assert np.array_equal(after["tr_q"][tr, :n], tr_q[tr, :n]), f"tree {tr}: configurations"
        assert np.array_equal(after["tr_parent"][tr, :n], tr_parent[tr, :n]), f"tree {tr}: parents"
        assert np.array_equal(after["tr_cost"][tr, :n], tr_cost[tr, :n]), f"tree {tr}: costs"
        assert np.array_equal(after["tr_rid"][tr, :n], tr_rid[tr, :n]), f"tree {tr}: roadmap ids"
    assert after["rng_round"] == before["rng_round"] + 1
    if winner is None:
        assert after["status"] == 0
    else:
        b, t, wnode, onode = winner
        assert after["status"] == b + 1, "solving expansion"
        path, cost = SR.record_path(tr_q, tr_parent, t, wnode, onode, ds.path_cap)
        assert after["path_n"] == len(path)
        assert np.array_equal(after["path"][:len(path)], path), "recorded path"
        # the path starts at the goal its chain reached (a root of tree 1) and ends at the start
        reached = [k for k, g in enumerate(p.goals) if np.array_equal(path[0], g)]
        assert reached and np.array_equal(path[-1], p.start)
        length = sum(float(np.linalg.norm(a.astype(np.float64) - q)) for a, q in zip(path[:-1], path[1:]))
        assert abs(length - cost) <= 1e-5 * length, "cost is the path's length (float32 sums in chain order)"
        # the best cost only falls (an optimise pass may record a dearer path)
        assert after["path_cost"] == cost and after["best"] == min(np.float32(best), cost)
    return dict(candidates=len(cand), edges=n_edges, inserted=len(index), connections=len(plans),
                rejected=rejected, rewired=len(rewires), solved=winner is not None,
                goal=reached[0] if winner is not None else None, exempt=exempt, n_commit=n_commit,
                inside=inside)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["wave", "split"])
@pytest.mark.parametrize("kind,seed", [("plain2", 4), ("table3", 2), ("bin4", 1), ("shelf5", 3)])
def test_search_rounds_match_the_restatement(ctx, kind, seed, form):
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    st = _lib.settings(num_new_configs=48, granularity=8, range=2.0)
    op = _lib.solve_options_default(block_per_robot=1, neighbour_candidates=8, neighbour_capacity=16,
                                    roadmap_capacity=4096, tree_capacity=4096)
    ds = _DiagSlot(ctx, p, st, op)
    try:
        # roadmap rounds until the start and first goal node of every arm
        # have a roadmap neighbour (else no expansion can move every arm)
        for _ in range(12):
            ds.roadmap(1)
            deg = ds.read("rm_deg", np.int32, (ds.A, ds.rm_cap))
            if (deg[:, :2] > 0).all():
                break
        stats = []
        keep = False  # the first round of a launch (round 0), then rounds continuing it
        with _expand_form(ctx, form):
            for r in range(6):
                before = ds.state()
                if before["status"] != 0:
                    break
                print(f"[search] {kind} seed {seed} {form} round {before['round']} trees {before['tr_n'].tolist()}",
                      flush=True)
                ds.search(1, keep)
                keep = True
                stats.append(_check_round(p, ds, before, ds.state(), st))
        print(f"[search] {kind} {form}: {stats}", flush=True)
        assert sum(x["edges"] for x in stats) > 0 and sum(x["inserted"] for x in stats) > 0
    finally:
        ds.close()


def _open_solved_slot(ctx, p, st, op, rounds=40):
    """A diagnostic slot whose first pass has run until solved."""
    ds = _DiagSlot(ctx, p, st, op)
    for _ in range(12):
        ds.roadmap(1)
        deg = ds.read("rm_deg", np.int32, (ds.A, ds.rm_cap))
        if (deg[:, :2] > 0).all():
            break
    keep = False
    for _ in range(rounds):
        ds.search(1, keep)
        keep = True
        if ds.read("counters", np.int32, (64,))[4] > 0:
            break
    return ds


def _check_simplify(p, ds, st):
    """Every smoothing stage and the shortcut stage of the slot's recorded
    path against oracle/simplify.py.  Returns (waypoint updates, exempt)."""
    from oracle import simplify as SP
    ctr = ds.read("counters", np.int32, (64,))
    N, D = int(ctr[7]), ds.D
    best = ctr[5:6].view(np.float32)[0]
    exempt = changed = 0
    stages = int(st.bspline_step)
    for stage in range(stages + 1):
        path = ds.read("path", np.float32, (ds.path_cap, D))[:N].copy()
        smooth = stage < stages
        ds.simplify(stage)
        fail0, mids, edges = SP.stage_edges(path, smooth, st.range, st.bspline_min_change, ds.sp_cap)
        n_items = N if smooth else N * 16
        assert int(ds.read("counters", np.int32, (64,))[12]) == len(edges)
        items = ds.read("sp_item", np.int32, (ds.sp_cap,))[:len(edges)]
        frm = ds.read("sp_from", np.float32, (ds.sp_cap, D))[:len(edges)]
        to = ds.read("sp_to", np.float32, (ds.sp_cap, D))[:len(edges)]
        coll = ds.read("sp_coll", np.uint32, (ds.sp_cap,))[:len(edges)] != 0
        assert items.tolist() == [e[0] for e in edges], "chunk order"
        if edges:
            assert np.array_equal(frm, np.array([e[1] for e in edges]).reshape(-1, D)), "chunk starts"
            assert np.array_equal(to, np.array([e[2] for e in edges]).reshape(-1, D)), "chunk ends"
            exempt += _verdicts_agree(p, O.MODE_FULL, ds.G, frm, to, coll)
        fail = fail0.copy()
        for it, c in zip(items, coll):
            if c:
                fail[it] = 1
        got_fail = ds.read("sp_fail", np.uint32, (ds.path_cap * 16,))[:n_items]
        assert np.array_equal(got_fail != 0, fail != 0), "candidate verdicts"
        after = ds.read("path", np.float32, (ds.path_cap, D))[:N]
        changed += int((after != path).any(axis=1).sum())
        if smooth:
            assert np.array_equal(after, SP.apply_smooth(path, fail, mids)), f"stage {stage}: smoothed path"
        else:
            want, cost, best2 = SP.shortcut(path, fail, best)
            assert np.array_equal(after, want), "shortcut path"
            c2 = ds.read("counters", np.int32, (64,))
            assert c2[6:7].view(np.float32)[0] == cost and c2[5:6].view(np.float32)[0] == best2
    return changed, exempt


_SMALL = dict(num_new_configs=48, granularity=8, range=2.0)
_SMALL_OPT = dict(block_per_robot=1, neighbour_candidates=8, neighbour_capacity=16, roadmap_capacity=4096,
                  tree_capacity=4096)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", [("table3", 2), ("shelf5", 3)])
def test_simplification_stages_match_the_restatement(ctx, kind, seed):
    """Path simplification (P9): every smoothing stage and the shortcut stage
    of a freshly recorded path against oracle/simplify.py -- candidates,
    chunk edges, the resulting path and its cost bit for bit; the chunk
    verdicts the device used against the FP64 oracle."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    st = _lib.settings(**_SMALL)
    ds = _open_solved_slot(ctx, p, st, _lib.solve_options_default(**_SMALL_OPT))
    try:
        assert ds.read("counters", np.int32, (64,))[4] > 0, "not solved in 40 rounds"
        changed, exempt = _check_simplify(p, ds, st)
        print(f"[simplify] {kind}: {changed} waypoint updates, {exempt} near-contact chunks", flush=True)
        assert changed > 0  # the stages did move waypoints
    finally:
        ds.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["plain2", "table3", "bin4", "shelf5"])
@pytest.mark.parametrize("seed", [10, 11, 12])
def test_rounds_and_simplification_sweep(ctx, kind, seed):
    """More problems through the same restatements: first-pass rounds until
    solved, the simplification stages of the recorded path, then optimise
    rounds (up to 30 rounds in all)."""
    _sweep(ctx, scenes.make_problem(kind, seed, _lib.gpu_checker(ctx)))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["plain2", "table3", "bin4", "shelf5"])
def test_rounds_and_simplification_sweep_five_launch_form(ctx, kind):
    """The same sweep with every expansion in the five-launch form."""
    with _expand_form(ctx, "split"):
        _sweep(ctx, scenes.make_problem(kind, 10, _lib.gpu_checker(ctx)))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", [("table3", 5), ("bin4", 6), ("shelf5", 7)])
def test_multi_goal_rounds_match_the_restatement(ctx, kind, seed):
    """Three goals: the goal tree starts with three roots, and a recorded
    path begins at the goal its chain reached."""
    _sweep(ctx, scenes.make_problem(kind, seed, _lib.gpu_checker(ctx), n_goals=3))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", [("table3", 2), ("shelf5", 3)])
def test_rounds_reaching_a_later_goal_match_the_restatement(ctx, kind, seed):
    """Goal 1 is a free waypoint halfway along a solved path (closer than
    goal 0): the rounds that solve record paths from the goal reached."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    st = _lib.settings(num_new_configs=64, granularity=16, range=2.0, path_simplify=0, rrtc_iter=1)
    sol = _lib.solve(ctx, p, st, _lib.solve_options_default(**_SMALL_OPT))
    assert sol.solved
    mid = interior_waypoint(sol.path)  # a tree node: the path is not simplified
    goals = [s["goal"] for s in _sweep(ctx, dataclasses.replace(p, goals=np.stack([p.goals[0], mid])))
             if s["goal"] is not None]
    print(f"[search] {kind}: goals reached {goals}", flush=True)
    assert 1 in goals


def _sweep(ctx, p):
    st = _lib.settings(num_new_configs=64, granularity=16, range=2.0)
    ds = _DiagSlot(ctx, p, st, _lib.solve_options_default(**_SMALL_OPT))
    try:
        for _ in range(12):
            ds.roadmap(1)
            deg = ds.read("rm_deg", np.int32, (ds.A, ds.rm_cap))
            if (deg[:, :2] > 0).all():
                break
        keep, opt, phs = False, False, None
        stats = []
        for _ in range(30):
            before = ds.state()
            if before["status"] != 0:
                if opt:
                    break
                _check_simplify(p, ds, st)
                ds.start_pass(1)
                opt, keep = True, False
                phs = (ds.read("phs_tf", np.float32, (ds.D, ds.D)), ds.read("phs_center", np.float32, (ds.D,)))
                before = ds.state()
            if not keep:
                before["round"] = 0
            ds.search(1, keep)
            keep = True
            stats.append(_check_round(p, ds, before, ds.state(), st, optimize=opt, phs=phs))
    finally:
        ds.close()
    return stats


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["wave", "split"])
@pytest.mark.parametrize("kind,seed", [("table3", 2), ("shelf5", 3)])
def test_optimise_pass_rounds_match_the_restatement(ctx, kind, seed, form):
    """An AO-x optimise pass (P3 PHS sampling, P4 cost-filtered nearest node,
    P5 bounded roadmap step, P6 rewire): after the first pass solves, the
    next pass starts with the best cost as bound and its rounds are restated
    from the device's samples (checked in float64 against the hyperspheroid
    sampler) -- rejections, cost bounds, candidates, rewires, inserts,
    connections and trees bit for bit."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    st = _lib.settings(**_SMALL)
    ds = _open_solved_slot(ctx, p, st, _lib.solve_options_default(**_SMALL_OPT))
    try:
        assert ds.read("counters", np.int32, (64,))[4] > 0, "first pass not solved in 40 rounds"
        ds.start_pass(1)
        phs = (ds.read("phs_tf", np.float32, (ds.D, ds.D)), ds.read("phs_center", np.float32, (ds.D,)))
        # the device samples with the host transform pinned in test_phs.py:
        # foci start and goal 0, transverse diameter the best cost
        best = ds.state()["best"]
        tf_ref, c_ref = test_phs.phs(p.start, p.goals[0], float(best))
        assert np.array_equal(phs[0], tf_ref.astype(np.float32)) and np.array_equal(phs[1], c_ref.astype(np.float32))
        cmin = float(np.linalg.norm(p.goals[0].astype(np.float64) - p.start))
        assert float(best) > cmin and np.linalg.norm(phs[0][:, 0]) == pytest.approx(float(best) / 2, rel=1e-5)
        stats = []
        keep = False
        with _expand_form(ctx, form):
            for _ in range(6):
                before = ds.state()
                if before["status"] != 0:
                    break
                if not keep:
                    before["round"] = 0  # a fresh launch starts its round count at 0
                ds.search(1, keep)
                keep = True
                stats.append(_check_round(p, ds, before, ds.state(), st, optimize=True, phs=phs))
        print(f"[optimise] {kind} {form}: {stats}", flush=True)
        # samples come from the informed set, so rejections are rare (clamping
        # to the joint box only); rewires must occur somewhere in the set
        assert sum(x["candidates"] for x in stats) > 0
        assert sum(x["inside"] for x in stats) > 0, "no hyperspheroid sample was checked"
        if kind == "table3":
            assert sum(x["rewired"] for x in stats) > 0
    finally:
        ds.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", [("bin4", 2), ("shelf5", 4)])
def test_whole_solves_agree_between_expansion_forms(ctx, kind, seed):
    """A whole solve (first pass, simplification, optimise passes), bounded by
    rounds rather than time, gives the same path and cost with every
    expansion in the wave form as in the five-launch form."""
    p = scenes.make_problem(kind, seed, _lib.gpu_checker(ctx))
    st = _lib.settings(num_new_configs=256, granularity=32, range=2.0, rrtc_iter=2, optimize_iters=300)
    op = _lib.solve_options_default(time_budget_s=120.0)
    sols = {}
    for form in ("wave", "split"):
        with _expand_form(ctx, form):
            print(f"[forms] {kind} seed {seed}: {form}", flush=True)
            sols[form] = _lib.solve(ctx, p, st, op)
    a, b = sols["wave"], sols["split"]
    assert a.solved and b.solved
    assert a.cost == b.cost and np.array_equal(a.path, b.path), "the forms disagree"
    assert a.info["rounds"] == b.info["rounds"] and a.info["passes"] == b.info["passes"]

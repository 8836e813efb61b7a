"""MR.POP on MI355X -- benchmark.

Default workload: the 4-arm bin-picking scene (4 Panda arms, 28 DoF, bin of 5
cuboid walls on a table plus 24 random spheres / cuboids / capsules), FP32,
granularity 32.  One step = one batched collision check of a fixed batch of
edges resident in HBM (FK + two-phase sphere checks of every interpolated
sample).  value = collision-checked edges per second on one GPU.

Also reported: the edge-check kernel's roofline (canonical FP32 work per
sample against the MI355X vector peak), the CPU baseline (the oracle's FP32
restatement of the reference checker, with the reference's early exits,
timed on a bounded sample of the same edges on this host's cores), and the
planner ("planning"): problems/s, time to first solution (p50/p99), solve
rate and median cost of the plain form (one problem at a time) against the
batched form (all problems in one engine), and of a serving run with Poisson
arrivals and per-request deadlines.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-plan]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path[:0] = [str(REPO), str(REPO / "mr.pop_amd")]

PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (published)
PEAK_HBM_GBS = 8000.0


def canonical_flops_per_sample(n_arms, n_cuboid, n_capsule, n_sphere):
    """W(A, O) = A*(942 + 10*f_O + 21*12) + C(A,2)*121*12 with
    f_O = 37*cuboids + 30*capsules + 12*spheres (coarse pass every sample runs)."""
    f_o = 37 * n_cuboid + 30 * n_capsule + 12 * n_sphere
    return n_arms * (942 + 10 * f_o + 21 * 12) + math.comb(n_arms, 2) * 121 * 12


def make_edges(problem, checker, n_edges, seed):
    from mrpop import scenes
    rng = np.random.default_rng(seed)
    pool = scenes.sample_free_configs(rng, problem.scene, problem.n_arms, problem.bases, 1024, checker)
    idx = rng.integers(len(pool), size=n_edges)
    frm = pool[idx]
    d = rng.normal(size=frm.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    to = (frm + d * rng.uniform(0.05, 1.0, size=(n_edges, 1))).astype(np.float32)
    return frm.astype(np.float32), to


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(problem, frm, to, G, budget_s=12.0):
    """Oracle FP32 restatement of the reference checker (early exits), timed
    on a bounded sample of the same edges: 1 thread, then all cores."""
    from oracle import oracle as O
    n = 256
    t0 = time.perf_counter()
    O.check_edges_f32(problem.scene, O.MODE_FULL, problem.bases, G, frm[:n], to[:n], threads=1)
    t1 = time.perf_counter() - t0
    rate1 = n / t1
    cores = os.cpu_count() or 1
    n_all = int(min(len(frm), max(n, rate1 * cores * budget_s / 2)))
    t0 = time.perf_counter()
    O.check_edges_f32(problem.scene, O.MODE_FULL, problem.bases, G, frm[:n_all], to[:n_all], threads=cores)
    tall = time.perf_counter() - t0
    return {"value": n_all / tall, "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": f"{n_all} edges of the timed batch (G={G}), all cores; 1 thread: {rate1:.0f} edges/s on {n} edges",
            "single_thread_value": rate1}


TRAFFIC_PROFILE = REPO / "profiles" / "check_kernel_bin4_g32_pmc.md"


def measured_traffic(n_edges, G):
    """HBM bytes per launch of the timed check (FETCH_SIZE + WRITE_SIZE from
    rocprofv3 --pmc, scaled by the counters' reading for a copy of known
    size), as recorded in the profile summary for this workload; None when
    the summary is absent or was taken on another batch."""
    if not TRAFFIC_PROFILE.exists():
        return None
    vals = {}
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("traffic:"):
            for kv in line[len("traffic:"):].split(","):
                k, v = kv.strip().split("=")
                vals[k] = float(v)
    if vals.get("edges") != n_edges or vals.get("granularity") != G:
        return None
    return vals.get("bytes_per_launch")


PLAN_KINDS = ("plain2", "table3", "bin4", "shelf5")   # 2 / 3 / 4 / 5 arms, ratio 1:1:1:1
SCENE_KINDS = ("bin4", "shelf5cell")   # 4 arms around a bin; 5 arms reaching into shelf cells
# per-problem time budget of the fixed-work scene lines: the 5-arm shelf-cell
# problems that go unsolved search the composite roadmap until the budget ends
SCENE_BUDGET_S = {"bin4": 30.0, "shelf5cell": 10.0}
# dRRT rounds of a launch before the next roadmap round while a problem has
# no solution yet (pop_settings::max_iters; the reference's default is 1000).
# Composite trees in the 5-arm shelf cells stop growing once the arms'
# roadmaps are exhausted around them, and the next roadmap round is what
# unblocks them: with 100, 15 of 16 shelf-cell problems are solved within
# 10 s against 10 of 16 with 1000.
SEARCH_ROUNDS = 100


def plan_problems(checker, n, first_seed=0):
    """n seeded problems, arm counts 2/3/4/5 in turn (seeds s, s, s, s, s + 1, ...)."""
    from mrpop import scenes
    return [scenes.make_problem(PLAN_KINDS[i % 4], first_seed + i // 4, checker) for i in range(n)]


def _stats(sols, ttfs, wall, budget_s=None):
    """Throughput, solve rate, median cost and TTFS percentiles of a run.
    ttfs_p50_s / ttfs_p99_s are over the solved problems ("ttfs_over");
    with a per-problem budget, ttfs_p99_at_budget_s counts every unsolved
    problem at the budget instead of leaving it out."""
    if not sols:
        return None
    solved = [s for s in sols if s.solved]
    t = np.array(ttfs, dtype=np.float64)
    costs = np.array([s.cost for s in solved]) if solved else np.array([np.nan])
    out = {"problems": len(sols), "wall_s": wall, "problems_per_s": len(sols) / wall,
           "solve_rate": len(solved) / len(sols), "unsolved": len(sols) - len(solved),
           "ttfs_over": "solved problems",
           "ttfs_p50_s": float(np.percentile(t, 50)) if len(t) else None,
           "ttfs_p99_s": float(np.percentile(t, 99)) if len(t) else None,
           "median_cost": float(np.median(costs)),
           "cost_p90": float(np.percentile(costs, 90)),
           "timeouts": sum(s.info["status"] == 4 for s in sols)}
    if budget_s is not None:
        full = np.concatenate([t, np.full(len(sols) - len(t), float(budget_s))])
        out["ttfs_p99_at_budget_s"] = float(np.percentile(full, 99))
    return out


def planning_bench(ctx, checker, n_plain, n_batch, plain_nnc, batch_nnc, serve_rates, serve_s, tune=True,
                   n_budget=0, batch_kc=16, batch_probe=0, n_scene=0, n_scene_plain=8, scene_kinds=SCENE_KINDS,
                   extras=True):
    """Plain vs batched planning on the same seeded problems, a budgeted
    batch, and serving."""
    from mrpop import _lib, autotune
    probs = plan_problems(checker, max(n_plain, n_batch))
    common = dict(granularity=32, range=2.0, rrtc_iter=6, optimize_iters=2000, max_iters=SEARCH_ROUNDS)
    op_common = dict(time_budget_s=30.0)
    st_plain = _lib.settings(num_new_configs=plain_nnc, **common)
    op = _lib.solve_options_default(**op_common)
    out = {"settings": dict(common, plain_num_new_configs=plain_nnc, time_budget_s=30.0),
           "problems": "seeded mixed 2/3/4/5-arm scenes (random cuboids, 3-arm table, 4-arm bin, 5-arm shelf), "
                       "one start and one goal each"}
    # batched-form settings: autotuned on a calibration set of other seeds
    tuned = {"num_new_configs": batch_nnc, "neighbour_candidates": batch_kc}
    if batch_probe:
        tuned["probe_samples"] = batch_probe
    if tune:
        calib = plan_problems(checker, max(16, n_batch), first_seed=1000)  # as many as the measured batch
        print(f"[bench] autotune on {len(calib)} calibration problems", file=sys.stderr, flush=True)
        res = autotune.tune(ctx, calib, common, dict(op_common, time_budget_s=5.0),
                            log=lambda m: print(f"[bench] autotune trial {m}", file=sys.stderr, flush=True))
        tuned = res.best
        out["autotune"] = res.as_dict()
        print(f"[bench] autotune: {tuned} in {res.seconds:.1f} s", file=sys.stderr, flush=True)
    tuned_st = {k: v for k, v in tuned.items() if k in autotune.SETTING_KEYS}
    op_kw = {k: v for k, v in tuned.items() if k not in autotune.SETTING_KEYS}
    st_batch = _lib.settings(**dict(common, **tuned_st))
    op_batch = _lib.solve_options_default(**op_common, **op_kw)
    out["settings"]["batched"] = tuned
    # plain: one problem at a time, all submitted at t = 0 (TTFS includes the queue)
    plain_sols, plain_ttfs, own_ttfs = [], [], []
    t0 = time.perf_counter()
    for p in probs[:n_plain]:
        start = time.perf_counter() - t0
        s = _lib.solve(ctx, p, st_plain, op)
        plain_sols.append(s)
        if s.solved:
            plain_ttfs.append(start + s.ttfs_s)
            own_ttfs.append(s.ttfs_s)
    plain = _stats(plain_sols, plain_ttfs, time.perf_counter() - t0) if plain_sols else None
    if plain is not None:
        plain["ttfs_own_p50_s"] = float(np.percentile(own_ttfs, 50)) if own_ttfs else None
        plain["ttfs_own_p99_s"] = float(np.percentile(own_ttfs, 99)) if own_ttfs else None
        print(f"[bench] plain: {plain}", file=sys.stderr, flush=True)
        out["plain"] = plain
    # batched: every problem in one engine, all submitted at t = 0
    if n_batch:
        t0 = time.perf_counter()
        batch_sols = _lib.solve_batch(ctx, probs[:n_batch], [st_batch] * n_batch, [op_batch] * n_batch)
        wall = time.perf_counter() - t0
        batched = _stats(batch_sols, [s.ttfs_s for s in batch_sols if s.solved], wall)
        print(f"[bench] batched: {batched}", file=sys.stderr, flush=True)
        out["batched"] = batched
        if plain is not None:
            both = [(a, b) for a, b in zip(plain_sols, batch_sols) if a.solved and b.solved]
            batched["median_cost_on_plain_problems"] = float(np.median([b.cost for _, b in both])) if both else None
            plain["median_cost_on_plain_problems"] = float(np.median([a.cost for a, _ in both])) if both else None
            out["batched_vs_plain_problems_per_s"] = batched["problems_per_s"] / plain["problems_per_s"]
    if n_budget:
        # the batched configuration with a fixed 2 s budget per problem and
        # the long optimisation passes of the 4-/5-arm configurations
        bprobs = plan_problems(checker, n_budget)
        st_b = _lib.settings(**dict(common, optimize_iters=400000, **tuned_st))
        op_b = _lib.solve_options_default(time_budget_s=2.0, **op_kw)
        t0 = time.perf_counter()
        sols = _lib.solve_batch(ctx, bprobs, [st_b] * n_budget, [op_b] * n_budget)
        out["batched_2s_budget"] = _stats(sols, [s.ttfs_s for s in sols if s.solved], time.perf_counter() - t0, 2.0)
        print(f"[bench] batched, 2 s budget: {out['batched_2s_budget']}", file=sys.stderr, flush=True)
    out["serving"] = [serving_bench(ctx, checker, rate, serve_s, st_batch, op_kw) for rate in serve_rates]
    if n_scene:
        out["scenes"] = scene_bench(ctx, checker, n_scene, max(1, n_scene_plain), op_kw, kinds=scene_kinds)
    if extras:
        out["two_arm_plain"] = two_arm_plain_bench(ctx, checker)
        out["same_scene"] = same_scene_bench()
    return out


def same_scene_bench(budget_s=30.0):
    """The reference's own evaluation: the 50 queries of each fixture scene
    (scripts/panda_four_problems.json, scripts/panda_five_problems.json --
    one scene per file) solved one after another in the plain form at its
    harness settings (nnc 300, G 64, range 2.0, rrtc_iter 3,
    scripts/evaluate_mr.cu:162-166; optimisation passes of 1500 rounds
    instead of 400 000 so that the 100 queries finish in seconds), without
    roadmap reuse (every query builds its roadmaps), and with roadmap_cache
    in one context: every query starts from the roadmaps the earlier queries
    of the scene left.  The batched form then plans the 50 queries of the
    scene together (fresh roadmaps, all submitted at t = 0); a query's result
    is the same alone and in the batch, so solve rate and costs match the
    fresh plain line and the batch changes throughput and TTFS only."""
    from mrpop import _lib, geometry, scenes
    st = _lib.settings(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=3, optimize_iters=1500)
    out = []
    for name, A in (("panda_four", 4), ("panda_five", 5)):
        probs = scenes.load_problem_file(REPO / "tests" / "golden" / f"{name}_problems.json",
                                         geometry.default_bases(A))
        res = {"scene": f"{name} fixture scene", "arms": A, "queries": len(probs),
               "settings": {"num_new_configs": 300, "granularity": 64, "range": 2.0, "rrtc_iter": 3,
                            "optimize_iters": 1500, "time_budget_s": budget_s}}
        fresh = None
        for cache in (0, 1):
            ctx = _lib.Context(0)
            op = _lib.solve_options_default(time_budget_s=budget_s, roadmap_cache=cache)
            sols, own, queued = [], [], []
            t0 = time.perf_counter()
            for p in probs:
                start = time.perf_counter() - t0
                s = _lib.solve(ctx, p, st, op)
                sols.append(s)
                if s.solved:
                    own.append(s.ttfs_s)
                    queued.append(start + s.ttfs_s)
            wall = time.perf_counter() - t0
            ctx.close()
            r = _stats(sols, own, wall, budget_s)
            r.update(_ttfs_pcts(own, "ttfs_own"), **_ttfs_pcts(queued, "ttfs_queued"))
            r["roadmap_rounds_mean"] = float(np.mean([s.info["roadmap_rounds"] for s in sols]))
            r["roadmap_nodes_built_mean"] = float(np.mean([sum(s.info["roadmap_size"][:A]) -
                                                           sum(s.info["roadmap_cached"][:A]) for s in sols]))
            res["roadmap_reuse" if cache else "fresh_roadmaps"] = r
            if not cache:
                fresh = sols
        ctx = _lib.Context(0)
        op = _lib.solve_options_default(time_budget_s=budget_s)
        t0 = time.perf_counter()
        bsols = _lib.solve_batch(ctx, probs, [st] * len(probs), [op] * len(probs))
        wall = time.perf_counter() - t0
        ctx.close()
        bt = [s.ttfs_s for s in bsols if s.solved]
        b = _stats(bsols, bt, wall, budget_s)
        b["same_cost_as_fresh_plain"] = sum(a.solved == c.solved and (not a.solved or a.cost == c.cost)
                                            for a, c in zip(fresh, bsols))
        res["batched"] = b
        res["batched_vs_plain_problems_per_s"] = b["problems_per_s"] / res["fresh_roadmaps"]["problems_per_s"]
        print(f"[bench] same scene: {res}", file=sys.stderr, flush=True)
        out.append(res)
    return out


def fixture_batch(ctx, name):
    """The 50 queries of one reference fixture scene in one batch at the
    harness settings (nnc 300, G 64, range 2.0, rrtc_iter 3, optimize_iters
    1500), after a batch of its first four queries (first launches)."""
    from mrpop import _lib, geometry, scenes
    A = 4 if name == "panda_four" else 5
    probs = scenes.load_problem_file(REPO / "tests" / "golden" / f"{name}_problems.json", geometry.default_bases(A))
    st = _lib.settings(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=3, optimize_iters=1500)
    op = _lib.solve_options_default(time_budget_s=30.0)
    _lib.solve_batch(ctx, probs[:4], [st] * 4, [op] * 4)
    t0 = time.perf_counter()
    sols = _lib.solve_batch(ctx, probs, [st] * len(probs), [op] * len(probs))
    wall = time.perf_counter() - t0
    out = _stats(sols, [s.ttfs_s for s in sols if s.solved], wall, 30.0)
    out.update(scene=f"{name} fixture scene", rounds=int(sum(s.info["rounds"] for s in sols)))
    return out


def _ttfs_pcts(ttfs, prefix):
    return {f"{prefix}_p50_s": float(np.percentile(ttfs, 50)) if len(ttfs) else None,
            f"{prefix}_p99_s": float(np.percentile(ttfs, 99)) if len(ttfs) else None}


def scene_bench(ctx, checker, n_problems, n_plain, option_kw, budget_s=2.0, kinds=SCENE_KINDS):
    """Per scene: n_problems seeded start/goal pairs at granularity 32,
    num_new_configs 1024, rrtc_iter 6, range 2.0.

    * fixed work (optimize_iters 2000, per-problem budget SCENE_BUDGET_S):
      the plain form (one problem at a time, all submitted at t = 0) on the
      first n_plain problems and the batched form (one engine, all
      submitted at t = 0) on all of them;
    * optimize_iters 400000 under a per-problem budget of budget_s, batched.

    TTFS percentiles are over the solved problems ("own": from the problem's
    own start; "queued": from t = 0)."""
    from mrpop import _lib, scenes
    out = []
    for kind in kinds:
        probs = [scenes.make_problem(kind, seed, checker) for seed in range(n_problems)]
        print(f"[bench] scene {kind}: {n_problems} problems generated", file=sys.stderr, flush=True)
        common = dict(granularity=32, range=2.0, rrtc_iter=6, num_new_configs=1024, max_iters=SEARCH_ROUNDS)
        budget = SCENE_BUDGET_S.get(kind, 30.0)
        st = _lib.settings(optimize_iters=2000, **common)
        op = _lib.solve_options_default(time_budget_s=budget, **option_kw)
        res = {"scene": kind, "arms": probs[0].n_arms, "problems": n_problems,
               "settings": dict(common, optimize_iters=2000, time_budget_s=budget, **option_kw)}
        sols, own, queued = [], [], []
        t0 = time.perf_counter()
        for p in probs[:n_plain]:
            start = time.perf_counter() - t0
            s = _lib.solve(ctx, p, st, op)
            sols.append(s)
            if s.solved:
                own.append(s.ttfs_s)
                queued.append(start + s.ttfs_s)
            print(f"[bench] scene {kind} plain {len(sols)}/{n_plain}: solved {s.solved} "
                  f"{time.perf_counter() - t0 - start:.2f} s", file=sys.stderr, flush=True)
        plain = _stats(sols, own, time.perf_counter() - t0, budget)
        plain.update(_ttfs_pcts(own, "ttfs_own"), **_ttfs_pcts(queued, "ttfs_queued"))
        t0 = time.perf_counter()
        bsols = _lib.solve_batch(ctx, probs, [st] * n_problems, [op] * n_problems)
        wall = time.perf_counter() - t0
        bt = [s.ttfs_s for s in bsols if s.solved]
        batched = _stats(bsols, bt, wall, budget)
        batched.update(_ttfs_pcts(bt, "ttfs_own"), **_ttfs_pcts(bt, "ttfs_queued"))
        both = [(a, b) for a, b in zip(sols, bsols) if a.solved and b.solved]
        plain["median_cost_on_common"] = float(np.median([a.cost for a, _ in both])) if both else None
        batched["median_cost_on_plain_problems"] = float(np.median([b.cost for _, b in both])) if both else None
        res.update(plain=plain, batched=batched,
                   batched_vs_plain_problems_per_s=batched["problems_per_s"] / plain["problems_per_s"])
        # the long optimisation of the scene's configuration under a fixed budget
        st_b = _lib.settings(optimize_iters=400000, **common)
        op_b = _lib.solve_options_default(time_budget_s=budget_s, **option_kw)
        t0 = time.perf_counter()
        s2 = _lib.solve_batch(ctx, probs, [st_b] * n_problems, [op_b] * n_problems)
        b2 = _stats(s2, [s.ttfs_s for s in s2 if s.solved], time.perf_counter() - t0, budget_s)
        b2["settings"] = dict(common, optimize_iters=400000, time_budget_s=budget_s)
        res["batched_budget"] = b2
        print(f"[bench] scene {kind}: {res}", file=sys.stderr, flush=True)
        out.append(res)
    return out


def two_arm_plain_bench(ctx, checker, budget_s=5.0):
    """The 2-arm cuboid scene (seed 0) in the plain form at granularity 32,
    num_new_configs 64, rrtc_iter 1 and range 0.5 rad, under a budget, and
    the same problem at range 2.0.  At range 0.5 a roadmap node of 7 joints
    has a fraction of a neighbour on average, so the arms' roadmaps do not
    connect start and goal within any practical budget."""
    from mrpop import _lib, scenes
    p = scenes.make_problem("plain2", 0, checker)
    out = {"scene": "plain2 (2 arms, 16 cuboids, seed 0)"}
    for rng_ in (0.5, 2.0):
        st = _lib.settings(granularity=32, range=rng_, num_new_configs=64, rrtc_iter=1)
        op = _lib.solve_options_default(time_budget_s=budget_s)
        t0 = time.perf_counter()
        s = _lib.solve(ctx, p, st, op)
        out[f"range_{rng_}"] = {"solved": s.solved, "cost": s.cost if s.solved else None,
                                "ttfs_s": s.ttfs_s if s.solved else None, "wall_s": time.perf_counter() - t0,
                                "status": s.info["status"], "roadmap_rounds": s.info["roadmap_rounds"],
                                "roadmap_size": s.info["roadmap_size"][:2], "budget_s": budget_s}
    print(f"[bench] 2-arm plain: {out}", file=sys.stderr, flush=True)
    return out


def serving_bench(ctx, checker, rate, duration_s, settings, option_kw, deadlines=(0.05, 0.25, 1.0), seed=7):
    """Poisson arrivals at `rate` req/s for `duration_s`; each request has a
    budget of the last deadline.  Reports TTFS from arrival and the cost of
    the best solution available at each deadline."""
    from mrpop import _lib
    rng = np.random.default_rng(seed)
    n = max(1, int(rate * duration_s))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    probs = plan_problems(checker, n, first_seed=2000)
    op = _lib.solve_options_default(time_budget_s=deadlines[-1], **option_kw)
    srv = _lib.Server(ctx, hbm_budget_bytes=200 * 10**9)
    tickets, done = [], {}
    t0 = time.perf_counter()
    i = 0
    while len(done) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            tickets.append(srv.submit(probs[i], settings, op))
            i += 1
        active, waiting = srv.step()
        if active == 0 and waiting == 0 and i < n:
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
        for k, t in enumerate(tickets):
            if t in done:
                continue
            state, sol = srv.poll(t, snapshot=False)
            if state == 2:
                done[t] = srv.poll(t, snapshot=True)[1]
    wall = time.perf_counter() - t0
    sols = [done[t] for t in tickets]
    ttfs = [s.ttfs_s for s in sols if s.solved]
    res = {"rate_req_s": rate, "requests": n, "wall_s": wall, "throughput_req_s": n / wall,
           "solve_rate": len(ttfs) / n,
           "ttfs_p50_s": float(np.percentile(ttfs, 50)) if ttfs else None,
           "ttfs_p99_s": float(np.percentile(ttfs, 99)) if ttfs else None}
    for d in deadlines:
        costs = []
        for s in sols:
            c = [cost for (cost, _), t in zip(s.seed_paths, s.improve_s) if t <= d]
            if c:
                costs.append(min(c))
        res[f"solved_by_{int(d * 1000)}ms"] = len(costs) / n
        res[f"median_cost_at_{int(d * 1000)}ms"] = float(np.median(costs)) if costs else None
    srv.close()
    print(f"[bench] serving: {res}", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--edges", type=int, default=65536)
    ap.add_argument("--granularity", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plan", action="store_true", help="skip the planner measurements")
    ap.add_argument("--plain-problems", type=int, default=8)
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--plain-nnc", type=int, default=1024)
    ap.add_argument("--batch-nnc", type=int, default=256)
    ap.add_argument("--serve-rates", type=str, default="20,50,100,200")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--budget-problems", type=int, default=256,
                    help="problems of the batched run with a 2 s budget each (0: skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched form with --batch-nnc / --batch-candidates instead of autotuning")
    ap.add_argument("--batch-candidates", type=int, default=16, help="roadmap candidates per node, untuned runs")
    ap.add_argument("--batch-probe", type=int, default=0, help="probe samples per planner edge, untuned runs")
    ap.add_argument("--plan-only", action="store_true", help="only the planner measurements (profiling runs)")
    ap.add_argument("--scene-problems", type=int, default=50,
                    help="seeded problems per scene of the 4-arm bin / 5-arm shelf-cell lines (0: skip)")
    ap.add_argument("--scene-plain", type=int, default=8, help="of which the plain form runs this many")
    ap.add_argument("--scenes", type=str, default=",".join(SCENE_KINDS), help="scenes of the per-scene lines")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the two-arm plain line and the same-scene fixture lines (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 256 MiB three times and exit: the reference kernel that scales the HBM counters")
    ap.add_argument("--fixture-batch", choices=("panda_four", "panda_five"),
                    help="plan the 50 queries of one reference fixture scene in one batch at the harness settings "
                         "and exit (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: problems are independent and this build runs on a single MI355X; "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    from mrpop import scenes, _lib

    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = 64 * 1024 * 1024  # floats: 256 MiB
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)
        return
    ctx = _lib.Context(0)
    if args.fixture_batch:
        print(json.dumps(fixture_batch(ctx, args.fixture_batch)), flush=True)
        return
    checker = _lib.gpu_checker(ctx)
    if args.plan_only:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        print(json.dumps(planning_bench(ctx, checker, args.plain_problems, args.batch_problems, args.plain_nnc,
                                        args.batch_nnc, rates, args.serve_seconds, not args.no_tune,
                                        args.budget_problems, args.batch_candidates, args.batch_probe,
                                        args.scene_problems, args.scene_plain, args.scenes.split(","),
                                        not args.no_extras)),
              flush=True)
        return
    problem = scenes.make_problem("bin4", 0, checker)
    G = args.granularity
    frm, to = make_edges(problem, checker, args.edges, seed=1)
    world = ctx.world(problem.scene, problem.n_arms, problem.bases)

    d_from = torch.from_numpy(frm).to(dev)
    d_to = torch.from_numpy(to).to(dev)
    d_out = torch.zeros(args.edges, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(dev)     # a real (non-null) stream: kernels and events share it
    sp = stream.cuda_stream

    def step():
        world.check_edges_device(_lib.MODE_FULL, G, args.edges, d_from.data_ptr(), d_to.data_ptr(),
                                 d_out.data_ptr(), None, sp)

    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_events = ev0.elapsed_time(ev1) / args.steps

    colliding = float(d_out.cpu().numpy().astype(bool).mean())
    # host-buffer entry point (copies over PCIe included): reported beside the value
    world.check_edges(_lib.MODE_FULL, G, frm, to)
    t_h = time.perf_counter()
    for _ in range(3):
        ref = world.check_edges(_lib.MODE_FULL, G, frm, to)
    host_edges_per_s = 3 * args.edges / (time.perf_counter() - t_h)
    assert np.array_equal(ref, d_out.cpu().numpy().astype(bool)), "device and host entry points disagree"

    edges_per_s = args.edges * args.steps / wall
    sc = problem.scene
    w_sample = canonical_flops_per_sample(problem.n_arms, len(sc.cuboids) + len(sc.z_cuboids),
                                          len(sc.capsules) + len(sc.z_capsules), len(sc.spheres))
    flops_per_launch = w_sample * G * args.edges
    achieved_tflops = flops_per_launch / (ms_events * 1e-3) / 1e12

    out = {
        "metric": "collision-checked edges/s; problems/s and p50/p99 TTFS, batched vs plain, 1 GPU",
        "value": edges_per_s,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 4-arm bin scene, edges from rejection-sampled free configurations",
        "config": {"workload": "bin4 edge check: 4 Panda arms (28 DoF), bin of 5 cuboids + table + 24 random "
                               "obstacles, granularity 32, batch of edges resident in HBM",
                   "edges_per_step": args.edges, "granularity": G, "arms": problem.n_arms,
                   "obstacles": sc.n_obstacles, "colliding_fraction": round(colliding, 4)},
        "host_buffers_edges_per_s": host_edges_per_s,
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": measured_traffic(args.edges, G),
                     "kernel": "check_kernel<4, FULL> (fused coarse + fine edge check)",
                     "ms_per_launch": ms_events, "flops_per_sample": w_sample},
    }
    print(f"[bench] edge check: {edges_per_s:.4g} edges/s, {ms_events:.3f} ms per launch", file=sys.stderr, flush=True)
    if not args.no_cpu:
        print("[bench] cpu baseline (bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(problem, frm, to, G)
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        # the 2-arm cuboid scene's edges: the CPU port on all cores beside this GPU
        p2 = scenes.make_problem("plain2", 0, checker)
        f2, t2 = make_edges(p2, checker, args.edges, seed=2)
        c2 = cpu_baseline(p2, f2, t2, G, budget_s=6.0)
        w2 = ctx.world(p2.scene, p2.n_arms, p2.bases)
        w2.check_edges(_lib.MODE_FULL, G, f2, t2)
        torch.cuda.synchronize()
        t_g = time.perf_counter()
        for _ in range(3):
            w2.check_edges(_lib.MODE_FULL, G, f2, t2)
        gpu2 = 3 * args.edges / (time.perf_counter() - t_g)
        out["cpu_baseline"]["two_arm_cuboids"] = {
            "cpu_edges_per_s": c2["value"], "cpu_single_thread_edges_per_s": c2["single_thread_value"],
            "cores": c2["cores"], "sample": c2["sample"], "gpu_edges_per_s_host_buffers": gpu2,
            "workload": "2 Panda arms (14 DoF), 16 cuboids, granularity 32, 65536 edges"}
    if not args.no_plan:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        out["planning"] = planning_bench(ctx, checker, args.plain_problems, args.batch_problems, args.plain_nnc,
                                         args.batch_nnc, rates, args.serve_seconds, not args.no_tune,
                                         args.budget_problems, args.batch_candidates, args.batch_probe,
                                         args.scene_problems, args.scene_plain, args.scenes.split(","),
                                         not args.no_extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

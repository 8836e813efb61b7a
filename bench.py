"""MR.POP on MI355X -- benchmark.

Default workload: the 4-arm bin-picking scene (4 Panda arms, 28 DoF, bin of 5
cuboid walls on a table plus 24 random spheres / cuboids / capsules), FP32,
granularity 32.  One step = one batched collision check of a fixed batch of
edges resident in HBM (FK + two-phase sphere checks of every interpolated
sample).  value = collision-checked edges per second on one GPU.

Also reported: the edge-check kernels' roofline (canonical FP32 work per
sample against the MI355X vector peak) and the CPU baseline (the oracle's
FP32 restatement of the reference checker, with the reference's early exits,
timed on a bounded sample of the same edges on this host's cores).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--edges", type=int, default=65536)
    ap.add_argument("--granularity", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: problems are independent and this build runs on a single MI355X; "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    from mrpop import scenes, _lib

    dev = torch.device("cuda:0")
    ctx = _lib.Context(0)
    checker = _lib.gpu_checker(ctx)
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
    ref = world.check_edges(_lib.MODE_FULL, G, frm, to)
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
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "coarse_kernel<4,FULL> + fine_kernel<4,FULL> (one edge check)",
                     "ms_per_launch": ms_events, "flops_per_sample": w_sample},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(problem, frm, to, G)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

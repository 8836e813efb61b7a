"""Command-line runs over problem files, with the outputs of the reference's
harness programs.

    python -m mrpop.evaluate evaluate <robot_name> <run_name> [options]
        One row per valid problem in <out>/<robot_name>_<run_name>.csv, with
        the header and row layout of scripts/evaluate_mr.cu:80-109 (settings
        num_new_configs 300, granularity 64, range 2.0, rrtc_iter 3, as in
        evaluate_mr.cu:163-166).

    python -m mrpop.evaluate export <robot_name> <out_dir> [rrtc_iter] [optimize_iters] [options]
        Seed paths as scripts/export_seeds.cu:89-160 writes them: <name>.txt
        (the final path) and <name>__kNN__cost<c>.txt for each improving
        solution, one waypoint per line, 9 significant digits.

robot_name is panda_four or panda_five (problem file
<problems_dir>/<robot_name>_problems.json, default tests/golden/), or any
problem file given with --problems.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

from . import _lib, geometry, scenes

REPO = Path(__file__).resolve().parents[2]
CSV_HEADER = ("solved,cost,path_length,start_tree_size,goal_tree_size,iters,wall_ns,kernel_ns,"
              "copy_ns,num_new_configs,granularity,range,balance,tree_ratio,dynamic_domain,dd_"
              "alpha,dd_radius,dd_min_radius,configs_str\n")


def fmt(x) -> str:
    """A float as a default-formatted C++ stream writes it (6 significant digits)."""
    return f"{float(x):g}"


def configs_str(path: np.ndarray) -> str:
    """print_cfg_to_ss (Planners.hh:43-49 of the reference): each value and a
    space, then a literal backslash-n after every configuration."""
    return "".join("".join(f"{fmt(v)} " for v in cfg) + "\\n" for cfg in path)


def csv_row(sol, st) -> str:
    i = sol.info
    fields = [str(int(sol.solved)), fmt(sol.cost), str(len(sol.path)), str(i["start_tree_size"]),
              str(i["goal_tree_size"]), str(i["iters"]), str(i["wall_ns"]), str(i["kernel_ns"]), str(i["copy_ns"]),
              str(st.num_new_configs), str(st.granularity), fmt(st.range), str(st.balance), fmt(st.tree_ratio),
              str(int(st.dynamic_domain)), fmt(st.dd_alpha), fmt(st.dd_radius), fmt(st.dd_min_radius)]
    return ", ".join(fields) + ", " + '"' + configs_str(sol.path) + '"' + "\n"


def path_text(path: np.ndarray) -> str:
    """write_path_file (export_seeds.cu:89-103): space-separated values, 9
    significant digits, one waypoint per line."""
    return "".join(" ".join(f"{float(v):.9g}" for v in cfg) + "\n" for cfg in path)


def seed_file_name(name: str, k: int, cost: float) -> str:
    return f"{name}__k{k:02d}__cost{cost:.4f}.txt"


def load(args):
    n_arms = {"panda_four": 4, "panda_five": 5}.get(args.robot_name)
    path = Path(args.problems) if args.problems else REPO / "tests" / "golden" / f"{args.robot_name}_problems.json"
    if n_arms is None and not args.problems:
        raise SystemExit(f"unsupported robot type: {args.robot_name}")
    probs = scenes.load_problem_file(path, geometry.default_bases(n_arms) if n_arms else None)
    return probs[:args.limit] if args.limit else probs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mrpop.evaluate", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    ev = sub.add_parser("evaluate")
    ev.add_argument("robot_name")
    ev.add_argument("run_name")
    ex = sub.add_parser("export")
    ex.add_argument("robot_name")
    ex.add_argument("out_dir")
    ex.add_argument("rrtc_iter", nargs="?", type=int, default=6)
    ex.add_argument("opt_iters", nargs="?", type=int, default=None, metavar="optimize_iters")
    for p in (ev, ex):
        p.add_argument("--problems", help="problem file (default: tests/golden/<robot_name>_problems.json)")
        p.add_argument("--out", default="test_output", help="CSV directory (evaluate)")
        p.add_argument("--limit", type=int, default=0, help="only the first N valid problems")
        p.add_argument("--optimize-iters", type=int, default=None, help="override optimize_iters")
        p.add_argument("--device", type=int, default=0)
        p.add_argument("--seed", type=int, default=1234, help="random stream of every problem")
    args = ap.parse_args(argv)

    ctx = _lib.Context(args.device)
    probs = load(args)
    kw = dict(num_new_configs=300, granularity=64, range=2.0)
    if args.cmd == "evaluate":
        kw["rrtc_iter"] = 3
    else:
        kw.update(rrtc_iter=args.rrtc_iter, phs=1, path_simplify=1)
    oi = args.optimize_iters
    if oi is None and args.cmd == "export":
        oi = args.opt_iters
    if oi is not None:
        kw["optimize_iters"] = oi
    st = _lib.settings(**kw)
    op = _lib.solve_options_default(time_budget_s=0)

    failed = 0
    if args.cmd == "evaluate":
        out_dir = Path(args.out)
        out_dir.mkdir(parents=True, exist_ok=True)
        out_path = out_dir / f"{args.robot_name}_{args.run_name}.csv"
        with open(out_path, "w") as f:
            f.write(CSV_HEADER)
            for p in probs:
                p.seed = args.seed
                sol = _lib.solve(ctx, p, st, op)
                failed += not sol.solved
                print(f"{p.name}: solved {int(sol.solved)} cost {fmt(sol.cost)} kernel (second): "
                      f"{sol.info['kernel_ns'] / 1e9:g}", flush=True)
                f.write(csv_row(sol, st))
        print(f"Total failed: {failed} / {len(probs)}  ({out_path})")
    else:
        out_dir = Path(args.out_dir)
        out_dir.mkdir(parents=True, exist_ok=True)
        for p in probs:
            p.seed = args.seed
            sol = _lib.solve(ctx, p, st, op)
            if not sol.solved:
                failed += 1
                print(f"[export] FAILED {p.name} (no seed written)")
                continue
            (out_dir / f"{p.name}.txt").write_text(path_text(sol.path))
            for k, (cost, path) in enumerate(sol.seed_paths):
                (out_dir / seed_file_name(p.name, k, cost)).write_text(path_text(path))
            print(f"[export] {p.name} -> {len(sol.path)} waypoints, cost {fmt(sol.cost)}, "
                  f"{len(sol.seed_paths)} intermediate seed(s)", flush=True)
        print(f"[export] wrote seeds for {len(probs) - failed}/{len(probs)} problems ({failed} failed) into {out_dir}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

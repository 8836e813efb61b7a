"""The drop-in C++ headers (mr.pop_amd/src/planning, mr.pop_amd/src/collision):
dRRT::solve<Robot> with the reference's signature (Planners.hh:57-64 of the
reference) through the C ABI.  The example program
mr.pop_amd/examples/solve_problem is built with the library."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from mrpop import _lib, geometry, scenes

REPO = Path(__file__).resolve().parents[1]
EXE = REPO / "mr.pop_amd" / "examples" / "solve_problem"


def _exe():
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    return str(EXE)


def test_headers_selftest():
    out = subprocess.run([_exe(), "--selftest"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "selftest: ok" in out.stdout


def test_bad_input_is_rejected(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("arms 9\n")
    out = subprocess.run([_exe(), str(p)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 2


def write_problem(path, problem, nnc, G, rng, rrtc_iter, optimize_iters):
    s = problem.scene
    lines = [f"arms {problem.n_arms}", f"settings {nnc} {G} {rng} {rrtc_iter} {optimize_iters}"]
    for key, arr in (("spheres", s.spheres), ("capsules", s.capsules), ("zcapsules", s.z_capsules),
                     ("cuboids", s.cuboids), ("zcuboids", s.z_cuboids)):
        lines.append(f"{key} {len(arr)}")
        lines += [" ".join(f"{v:.9g}" for v in rec) for rec in arr]
    lines.append("start " + " ".join(f"{v:.9g}" for v in problem.start))
    lines.append(f"goals {len(problem.goals)}")
    lines += [" ".join(f"{v:.9g}" for v in g) for g in problem.goals]
    path.write_text("\n".join(lines) + "\n")


def parse_output(text, dim):
    lines = text.strip().splitlines()
    solved = int(lines[0].split()[1])
    cost = float(lines[1].split()[1])
    n = int(lines[2].split()[1])
    path = np.array([[float(v) for v in ln.split()] for ln in lines[3:3 + n]], np.float32).reshape(n, dim)
    return solved, cost, path


@pytest.mark.gpu
def test_drop_in_solve_matches_the_c_abi(tmp_path, ctx):
    """dRRT::solve<Panda_four> on a reference fixture problem returns the
    path that mrpop_solve returns for the same problem, settings and seed
    (1234, the reference's generator seed)."""
    import dataclasses
    from test_gpu_plan import check_solution
    probs = scenes.load_problem_file(str(REPO / "tests/golden/panda_four_problems.json"), geometry.default_bases(4))
    p = dataclasses.replace(probs[0], seed=1234)
    nnc, G, rng, rrtc, oi = 300, 64, 2.0, 2, 500
    f = tmp_path / "p.txt"
    write_problem(f, p, nnc, G, rng, rrtc, oi)
    print("[dropin] solve_problem", p.name, flush=True)
    out = subprocess.run([_exe(), str(f)], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    solved, cost, path = parse_output(out.stdout, p.dim)
    assert solved == 1
    st = _lib.settings(num_new_configs=nnc, granularity=G, range=rng, rrtc_iter=rrtc, optimize_iters=oi)
    ref = _lib.solve(ctx, p, st, _lib.solve_options_default(time_budget_s=0))
    assert np.array_equal(ref.path, path)
    assert cost == pytest.approx(ref.cost, rel=1e-6)
    check_solution(p, ref, granularity=G, rng=rng)


@pytest.mark.gpu
def test_drop_in_reports_rejected_settings(tmp_path):
    """Settings the planner rejects reach the C++ caller as an unsolved
    result with the reason printed, as the drop-in's other errors are."""
    probs = scenes.load_problem_file(str(REPO / "tests/golden/panda_four_problems.json"), geometry.default_bases(4))
    f = tmp_path / "p.txt"
    write_problem(f, probs[0], 300, 0, 2.0, 2, 500)  # granularity 0
    print("[dropin] solve_problem granularity 0", flush=True)
    out = subprocess.run([_exe(), str(f)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 3, out.stdout[-2000:] + out.stderr[-2000:]
    assert "granularity outside 1..64" in out.stderr
    assert parse_output(out.stdout, probs[0].dim)[0] == 0

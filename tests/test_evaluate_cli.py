"""Output formats of python -m mrpop.evaluate: the CSV rows of the
reference's scripts/evaluate_mr.cu:80-109 and the seed files of
scripts/export_seeds.cu:89-160."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from mrpop import _lib, evaluate

REPO = Path(__file__).resolve().parents[1]


class _Sol:
    solved = True
    cost = float(np.float32(13.63115))  # costs come from float32
    path = np.array([[0.1, -2.0, 3.25], [1e-5, 0.0, 123456.7]], np.float32)
    info = dict(start_tree_size=10, goal_tree_size=20, iters=7, wall_ns=123, kernel_ns=45, copy_ns=6)


def test_csv_header_and_row_layout():
    st = _lib.settings(num_new_configs=300, granularity=64, range=2.0, rrtc_iter=3)
    assert evaluate.CSV_HEADER.startswith("solved,cost,path_length,start_tree_size,goal_tree_size,iters,wall_ns")
    assert evaluate.CSV_HEADER.rstrip("\n").endswith("dd_alpha,dd_radius,dd_min_radius,configs_str")
    row = evaluate.csv_row(_Sol(), st)
    fields = row.rstrip("\n").split(", ")
    assert fields[:9] == ["1", "13.6312", "2", "10", "20", "7", "123", "45", "6"]
    assert fields[9:18] == ["300", "64", "2", "2", "1", "0", "0.0001", "4", "1"]
    # configs_str: "%g " per value, a literal backslash-n after each configuration
    assert fields[18] == '"0.1 -2 3.25 \\n1e-05 0 123457 \\n"'


def test_seed_file_names_and_text():
    assert evaluate.seed_file_name("bin_test0", 3, 14.56789) == "bin_test0__k03__cost14.5679.txt"
    txt = evaluate.path_text(np.array([[0.123456789012, -1.0]], np.float32))
    assert txt == f"{np.float32(0.123456789012):.9g} -1\n"


def test_cli_usage():
    out = subprocess.run([sys.executable, "-m", "mrpop.evaluate", "--help"], capture_output=True, text=True,
                         cwd=REPO / "mr.pop_amd", timeout=120)
    assert out.returncode == 0 and "evaluate" in out.stdout and "export" in out.stdout


@pytest.mark.gpu
def test_cli_runs_fixture_problems(tmp_path):
    env = dict(__import__("os").environ, PYTHONPATH=f"{REPO}:{REPO / 'mr.pop_amd'}")
    print("[cli] evaluate panda_four", flush=True)
    out = subprocess.run([sys.executable, "-m", "mrpop.evaluate", "evaluate", "panda_four", "t", "--limit", "1",
                          "--optimize-iters", "300", "--out", str(tmp_path)], capture_output=True, text=True,
                         env=env, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    rows = (tmp_path / "panda_four_t.csv").read_text().splitlines()
    assert rows[0] + "\n" == evaluate.CSV_HEADER and len(rows) == 2 and rows[1].startswith("1, ")
    print("[cli] export panda_five", flush=True)
    out = subprocess.run([sys.executable, "-m", "mrpop.evaluate", "export", "panda_five", str(tmp_path / "seeds"),
                          "2", "300", "--limit", "1"], capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    files = sorted(p.name for p in (tmp_path / "seeds").iterdir())
    assert any("__k00__cost" in f for f in files) and any("__" not in f for f in files)
    final = [f for f in files if "__" not in f][0]
    rows = (tmp_path / "seeds" / final).read_text().splitlines()
    assert all(len(r.split()) == 35 for r in rows) and len(rows) >= 2

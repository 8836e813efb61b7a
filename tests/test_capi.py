"""CPU tests of the C ABI: the library loads, exports every function that
include/mrpop_c.h declares, and its host-only entry points behave (no GPU
is touched here)."""
import ctypes
import json
import re

import numpy as np
import pytest

from conftest import GOLDEN, REPO
from mrpop import _lib, geometry


def declared_functions():
    text = _lib.HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mrpop_\w+)\s*\(", text)))


def test_header_functions_are_exported_and_bound():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/mrpop_c.h but not exported"
        assert n in _lib.SIGNATURES, f"{n} not bound in mrpop/_lib.py"
    assert set(_lib.SIGNATURES) <= set(names)


def test_settings_defaults_match_pop_settings():
    s = _lib.settings_default()
    # src/planning/pop_settings.hh:4-31
    want = dict(max_samples=1000000, max_iters=1000, optimize_iters=400000, num_new_configs=256, granularity=16,
                range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, dd_alpha=0.0001, dd_radius=4.0,
                dd_min_radius=1.0, roadmap_reassign=0, rrtc_iter=1, resample_iter=5, cost_bound_resample=0,
                path_simplify=1, phs=1, simplify_round=1, bspline_step=3, bspline_min_change=0.01, perturb_step=3,
                perturb_empty_step=2)
    for k, v in want.items():
        assert getattr(s, k) == pytest.approx(v, rel=1e-6), k


def test_solve_options_defaults():
    """Solve options of this library: the roadmap steering period is the
    reference's BIASED_ITER (roadmap_interior.cuh:24), roadmap reuse across
    queries is off, block_per_robot is the harness's 8 and the roadmap
    capacity MAX_ROADMAP_SIZE (roadmap_interior.cuh:21)."""
    o = _lib.solve_options_default()
    assert o.biased_iter == 10000 and o.roadmap_cache == 0
    assert o.block_per_robot == 8 and o.roadmap_capacity == 100000
    assert o.graph_rounds == 0 and o.probe_samples == 0


def test_default_bases():
    tables = json.loads((GOLDEN / "panda_tables.json").read_text())
    for A in (2, 3, 4, 5):
        out = np.zeros((A, 12), np.float32)
        assert _lib.lib().mrpop_default_bases(A, out.ctypes.data) == _lib.OK
        assert np.array_equal(out, geometry.default_bases(A))
        if A >= 4:
            ref = np.array(tables[f"base_{A}"], np.float32).reshape(A, 4, 4)[:, :3].reshape(A, 12)
            assert np.array_equal(out, ref)
        R = out.reshape(A, 3, 4)[:, :, :3].astype(np.float64)
        assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-6)
    assert _lib.lib().mrpop_default_bases(6, np.zeros(72, np.float32).ctypes.data) == _lib.ERR_INVALID


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.mrpop_settings_default(None) == _lib.ERR_INVALID
    assert L.mrpop_create(0, None) == _lib.ERR_INVALID
    assert L.mrpop_destroy(None) == _lib.OK
    assert L.mrpop_last_error(None) == b"null context"
    assert L.mrpop_version().startswith(b"mrpop")


def test_version_names_target():
    assert b"gfx950" in _lib.lib().mrpop_version()

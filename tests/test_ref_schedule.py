"""The reference's fine-stage thread schedule and what it changes.

The reference gives each of a block's 4*G threads to one sample of the edge
in proportion to the sample's fine sphere count (utils.cuh:477-517,
653-676); a sample whose share rounds down to zero has no thread, so its fine
stage never runs and a collision there is missed.  This build checks every
sample.  The oracle restates the schedule (orc_check_edges_ref_schedule);
these tests pin it by hand-computed shares and measure how often it changes
an edge verdict.
"""
import numpy as np
import pytest

from mrpop import geometry, scenes
from oracle import oracle as O
from conftest import GOLDEN


@pytest.mark.parametrize("counts, shares", [
    # 4 samples, 16 threads: 27/30*16 = 14.4 -> 14, 1/30*16 = 0.53 -> 0, and
    # the 2 threads the floors leave go to the last sample
    ([27, 1, 1, 1], [14, 0, 0, 2]),
    # 8 threads: 1/101*8 = 0.079 -> 0, 100/101*8 = 7.92 -> 7, leftover 1
    ([1, 100], [0, 8]),
    # one flagged sample takes every thread
    ([0, 0, 5, 0], [0, 0, 16, 0]),
    # equal counts split evenly: 12/48*32 = 8 each
    ([12, 12, 12, 12, 0, 0, 0, 0], [8, 8, 8, 8, 0, 0, 0, 0]),
    # nothing flagged: the stage does not run; the leftover sits on the last sample
    ([0, 0, 0], [0, 0, 12]),
])
def test_thread_shares_known_answers(counts, shares):
    got = O.ref_thread_shares(counts)
    assert got.tolist() == shares
    assert got.sum() == 4 * len(counts)


def test_shares_cover_every_thread_once():
    """Every thread t < 4G lands on exactly one sample (the first whose
    running share total exceeds t), and a sample has a thread iff its share
    is positive."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        G = int(rng.integers(1, 65))
        counts = rng.integers(0, 120, size=G) * (rng.random(G) < 0.3)
        sh = O.ref_thread_shares(counts)
        if counts.sum() == 0:
            continue
        owner = [int(np.argmax(np.cumsum(sh) - 1 >= t)) for t in range(4 * G)]
        assert len(owner) == 4 * G
        has = np.zeros(G, bool)
        has[owner] = True
        assert np.array_equal(has, sh > 0)
        # samples without flagged spheres never get a thread, except the last
        assert not (sh[:-1][counts[:-1] == 0] > 0).any()


def _golden():
    for f in sorted(GOLDEN.glob("edges_*.npz")):
        z = np.load(f)
        yield f.name, z, scenes.Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])


def test_schedule_only_drops_checks_on_golden_edges():
    """Under the schedule an edge can only turn from colliding to free; on
    the golden edges (13 files, 2 912 edges, 1 186 colliding) it turns none:
    the few colliding samples left without a thread sit on edges that
    collide elsewhere too."""
    dropped_edges = 0
    for name, z, sc in _golden():
        r = O.check_edges_ref_schedule(sc, int(z["mode"]), z["bases"], int(z["G"]), z["frm"], z["to"],
                                       arm=int(z["arm"]))
        assert np.array_equal(r["all"], z["mask_nom"] != 0), name
        assert not (r["sched"] & ~r["all"]).any(), name
        assert np.array_equal(r["sched"], r["all"]), name
        assert ((r["dropped"] & ~z["mask_nom"]) == 0).all(), name
        dropped_edges += int((r["dropped"] != 0).sum())
    assert dropped_edges > 0   # the schedule does leave colliding samples unchecked


@pytest.mark.parametrize("path, arms", [("panda_four_problems.json", 4), ("panda_five_problems.json", 5)])
def test_schedule_on_planner_steps(path, arms):
    """Steps of 0.25-2 rad per arm from free configurations of the reference's
    fixture scenes, at the planner's FULL (G = 32) and DRRT (G = 16)
    resolutions: no edge verdict changes under the schedule."""
    rng = np.random.default_rng(11)
    probs = scenes.load_problem_file(GOLDEN / path, geometry.default_bases(arms))
    colliding = 0
    for p in probs[:3]:
        cfg = scenes.sample_free_configs(rng, p.scene, arms, p.bases, 64, O.config_checker)
        for step in (0.25, 2.0):
            d = rng.normal(size=(len(cfg), arms, 7))
            d /= np.linalg.norm(d, axis=2, keepdims=True)
            to = (cfg.reshape(len(cfg), arms, 7) + step * d).reshape(len(cfg), -1).astype(np.float32)
            for mode, G in ((O.MODE_FULL, 32), (O.MODE_DRRT, 16)):
                r = O.check_edges_ref_schedule(p.scene, mode, p.bases, G, cfg, to)
                assert not (r["sched"] & ~r["all"]).any()
                assert np.array_equal(r["sched"], r["all"])
                colliding += int(r["all"].sum())
    assert colliding > 100

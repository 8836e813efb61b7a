"""Write seeded edge batches with FP64 oracle verdicts (edges_*.npz).

Each file holds: scene records, bases, mode, arm, granularity, from/to
(float32), and the oracle's per-edge sample masks under three readings of
near-contact tests (|clearance| < 1e-5 m): nominal, as free (lo) and as
collision (hi).  Samples where lo and hi differ are exempt from bit parity.

Edges start at collision-free configurations (the reference fixtures'
start/goal configurations, or rejection-sampled ones) and step 0.02-1.5 rad in
a random direction, so the batches mix free, grazing and colliding edges.

    python tests/golden/gen_edge_vectors.py               # the original sets
    python tests/golden/gen_edge_vectors.py zaligned quirk  # the later sets
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path[:0] = [str(REPO), str(REPO / "mr.pop_amd")]

from oracle import oracle as O  # noqa: E402
from mrpop import scenes, geometry  # noqa: E402


def edges_from(rng, starts, n, dim, lo=0.02, hi=1.5):
    idx = rng.integers(len(starts), size=n)
    frm = starts[idx].astype(np.float32)
    d = rng.normal(size=(n, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    step = rng.uniform(lo, hi, size=(n, 1))
    to = (frm + d * step).astype(np.float32)
    return frm, to


def save(name, scene, bases, mode, arm, G, frm, to):
    r = O.check_edges(scene, mode, bases, G, frm, to, arm=arm)
    np.savez_compressed(HERE / f"edges_{name}.npz", spheres=scene.spheres, capsules=scene.capsules,
                        z_capsules=scene.z_capsules, cuboids=scene.cuboids, z_cuboids=scene.z_cuboids,
                        bases=np.asarray(bases, np.float32), mode=mode, arm=arm, G=G, frm=frm, to=to,
                        mask_nom=r["mask_nom"], mask_lo=r["mask_lo"], mask_hi=r["mask_hi"])
    print(name, "edges", len(frm), "colliding", r["nom"].mean().round(3), "exempt edges", int((r["lo"] != r["hi"]).sum()))


def zaligned_sets():
    """A 3-arm cell of z-aligned posts (z-capsules) and yaw-rotated boxes
    (z-cuboids), the obstacle types of utils.cuh:165-181 and 208-228, in the
    full multi-arm mode and for one arm with its base pose."""
    rng = np.random.default_rng(20261204)
    p = scenes.make_problem("posts3", 7, O.config_checker, n_goals=15)
    assert len(p.scene.z_capsules) and len(p.scene.z_cuboids)
    starts = np.concatenate([p.start[None], p.goals])
    frm, to = edges_from(rng, starts, 256, p.dim)
    save("posts3_full_g32", p.scene, p.bases, O.MODE_FULL, 0, 32, frm, to)
    frm, to = edges_from(rng, starts[:, 14:21], 256, 7)
    save("posts3_single_g64", p.scene, p.bases, O.MODE_SINGLE, 2, 64, frm, to)


def quirk_set():
    """Edges of the bin-picking fixture scene (DRRT mode) that end at
    configurations where the reference's multi-arm reading of the fine self
    ranges (panda_four.cuh:390-401) gives another verdict than the
    single-arm reading; mask_plain holds that other reading."""
    rng = np.random.default_rng(20261205)
    probs = scenes.load_problem_file(HERE / "panda_four_problems.json", geometry.default_bases(4))
    scene, bases = probs[0].scene, probs[0].bases
    starts = np.concatenate([np.stack([p.start for p in probs]), np.concatenate([p.goals for p in probs])])
    low, scale = np.tile(scenes.LOW, 4), np.tile(scenes.SCALE, 4)
    ends = []
    while len(ends) < 96:
        q = (rng.uniform(0, 1, (4096, 28)) * scale + low).astype(np.float32)
        a = O.check_edges(scene, O.MODE_DRRT, bases, 1, q, q)
        b = O.check_edges_plain_ranges(scene, O.MODE_DRRT, bases, 1, q, q)
        robust = (a["lo"] == a["hi"]) & (b["lo"] == b["hi"])
        ends.extend(q[robust & (a["nom"] != b["nom"])])
    end = np.stack(ends[:96])
    frm = starts[rng.integers(len(starts), size=len(end))].astype(np.float32)
    save("panda_four_quirk_drrt_g16", scene, bases, O.MODE_DRRT, 0, 16, frm, end)
    r = O.check_edges_plain_ranges(scene, O.MODE_DRRT, bases, 16, frm, end)
    z = dict(np.load(HERE / "edges_panda_four_quirk_drrt_g16.npz"))
    z["mask_plain"] = r["mask_nom"]
    np.savez_compressed(HERE / "edges_panda_four_quirk_drrt_g16.npz", **z)


EXTRA = {"zaligned": zaligned_sets, "quirk": quirk_set}


def main():
    if len(sys.argv) > 1:  # named extra sets only, e.g. `gen_edge_vectors.py zaligned quirk`
        for name in sys.argv[1:]:
            EXTRA[name]()
        return
    rng = np.random.default_rng(20261203)
    for fname, A in (("panda_four", 4), ("panda_five", 5)):
        probs = scenes.load_problem_file(HERE / f"{fname}_problems.json", geometry.default_bases(A))
        scene, bases = probs[0].scene, probs[0].bases
        starts = np.concatenate([np.stack([p.start for p in probs]), np.concatenate([p.goals for p in probs])])
        frm, to = edges_from(rng, starts, 256, 7 * A)
        save(f"{fname}_full_g32", scene, bases, O.MODE_FULL, 0, 32, frm, to)
        frm, to = edges_from(rng, starts, 256, 7 * A)
        save(f"{fname}_drrt_g16", scene, bases, O.MODE_DRRT, 0, 16, frm, to)
        # single arm (roadmap check): arm 1's slice, G = 64
        frm, to = edges_from(rng, starts[:, 7:14], 256, 7)
        save(f"{fname}_single_g64", scene, bases, O.MODE_SINGLE, 1, 64, frm, to)
    # synthetic scenes with spheres, capsules and rotated cuboids
    for kind in ("plain2", "table3", "bin4", "shelf5"):
        p = scenes.make_problem(kind, 7, O.config_checker, n_goals=15)
        starts = np.concatenate([p.start[None], p.goals])
        frm, to = edges_from(rng, starts, 192, p.dim)
        save(f"{kind}_full_g32", p.scene, p.bases, O.MODE_FULL, 0, 32, frm, to)


if __name__ == "__main__":
    main()

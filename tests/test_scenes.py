"""Scene generators of the benchmark configurations (CPU): the hand
positions used to place end-effectors in shelf cells against the FP64
oracle's sphere centres, and the properties the shelf-cell and z-obstacle
problems promise."""
import numpy as np

from mrpop import geometry, scenes
from oracle import oracle as O


def test_hand_positions_match_the_oracle():
    rng = np.random.default_rng(3)
    bases = geometry.default_bases(5)
    q = rng.uniform(np.tile(scenes.LOW, 5), np.tile(scenes.LOW + scenes.SCALE, 5), (40, 35)).astype(np.float32)
    for a in range(5):
        got = scenes.hand_positions(bases[a], q[:, 7 * a:7 * a + 7])
        ref = np.array([np.asarray(O.sphere_centres(bases, qq)[0])[a][7:11].mean(axis=0) for qq in q])
        assert np.abs(got - ref).max() < 1e-9


def test_shelf_cell_problems():
    """Every hand inside a shelf cell, the arms in distinct cells, at the
    start and at each goal, and every configuration collision-free."""
    cells = scenes.shelf_cells()
    for seed in (0, 1):
        p = scenes.make_problem("shelf5cell", seed, O.config_checker, n_goals=2)
        cfgs = np.vstack([p.start[None], p.goals])
        assert not O.config_checker(p.scene, 5, p.bases, cfgs).any()
        for cfg in cfgs:
            where = []
            for a in range(5):
                h = scenes.hand_positions(p.bases[a], cfg[7 * a:7 * a + 7])[0]
                inside = [k for k, (lo, hi) in enumerate(cells) if (h > lo).all() and (h < hi).all()]
                assert len(inside) == 1, f"arm {a}: hand not in a cell"
                where.append(inside[0])
            assert len(set(where)) == 5, "two hands share a cell"


def test_posts_scene_uses_z_aligned_obstacles():
    p = scenes.make_problem("posts3", 2, O.config_checker, n_goals=2)
    assert len(p.scene.z_capsules) == 10 and len(p.scene.z_cuboids) == 10
    zc = p.scene.z_capsules
    assert np.allclose(zc[:, 3:5], 0) and np.allclose(zc[:, 7], 1.0 / zc[:, 5] ** 2, rtol=1e-5)
    zb = p.scene.z_cuboids
    assert np.allclose(zb[:, [5, 8, 9, 10]], 0) and np.allclose(zb[:, 11], 1)
    assert np.allclose(zb[:, 3] ** 2 + zb[:, 4] ** 2, 1, atol=1e-6)

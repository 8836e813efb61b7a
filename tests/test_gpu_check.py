"""GPU parity of the edge checker (libmrpop.so through the C ABI) against the
FP64 oracle and the committed golden vectors.

Bar: per-sample verdicts bit-identical to the oracle, except samples whose
oracle verdict flips when tests within 1e-5 m of contact are read as free
or as colliding (mask_lo != mask_hi).  FK sphere centres within 1e-5 m of
the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FK_TOL = 1e-5  # metres


def _check_probe(w, mode, G, frm, to, probe, arm=0):
    """Edge verdicts of the two-pass schedule the planner uses: `probe`
    samples per edge, then every sample of the edges still free."""
    import ctypes
    from mrpop._lib import lib
    f = lib().mrpop_diag_check_probe
    f.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2 + [ctypes.c_int, ctypes.c_void_p]
    frm = np.ascontiguousarray(frm, np.float32)
    to = np.ascontiguousarray(to, np.float32)
    out = np.zeros(len(frm), np.uint8)
    w.ctx.check(f(w.ctx.handle, w._h, mode, arm, G, len(frm), frm.ctypes.data, to.ctypes.data, probe,
                  out.ctypes.data))
    return out.astype(bool)


def _scene_from_npz(z):
    from mrpop.scenes import Scene
    return Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])


def _assert_masks(gpu_mask, gpu_edge, mask_lo, mask_hi, label):
    gpu_mask = gpu_mask.astype(np.uint64)
    exempt = mask_lo ^ mask_hi                       # samples a 1e-5 m change could flip
    bad = (gpu_mask ^ mask_lo) & ~exempt
    n_bad = int(np.count_nonzero(bad))
    assert n_bad == 0, f"{label}: {n_bad} edges with sample verdicts differing from the oracle"
    edge_robust = (mask_lo != 0) == (mask_hi != 0)
    assert np.array_equal(gpu_edge[edge_robust], (mask_lo != 0)[edge_robust]), f"{label}: edge verdicts"
    assert np.array_equal(gpu_edge, gpu_mask != 0), f"{label}: edge verdict != OR of sample mask"


GOLDEN_FILES = sorted(p.name for p in GOLDEN.glob("edges_*.npz"))


@pytest.mark.parametrize("name", GOLDEN_FILES)
def test_golden_edges(ctx, name):
    z = np.load(GOLDEN / name)
    scene = _scene_from_npz(z)
    bases = z["bases"]
    mode, arm, G = int(z["mode"]), int(z["arm"]), int(z["G"])
    print(f"[gpu] golden {name} mode={mode} G={G} edges={len(z['frm'])}", flush=True)
    w = ctx.world(scene, len(bases), bases)
    edge, mask = w.check_edges(mode, G, z["frm"], z["to"], arm=arm, with_mask=True)
    _assert_masks(mask, edge, z["mask_lo"], z["mask_hi"], name)
    # without the mask output the kernel may skip samples of known-colliding edges
    edge2 = w.check_edges(mode, G, z["frm"], z["to"], arm=arm)
    assert np.array_equal(edge2, edge)
    for probe in (1, 4, 7):
        assert np.array_equal(_check_probe(w, mode, G, z["frm"], z["to"], probe, arm=arm), edge), probe


def _random_edges(rng, starts, n, dim, lo, hi):
    idx = rng.integers(len(starts), size=n)
    frm = starts[idx].astype(np.float32)
    d = rng.normal(size=(n, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    to = (frm + d * rng.uniform(lo, hi, size=(n, 1))).astype(np.float32)
    return frm, to


@pytest.mark.parametrize("kind", ["plain2", "table3", "bin4", "shelf5", "posts3"])
@pytest.mark.parametrize("G", [1, 16, 32, 64])
def test_live_oracle_parity(ctx, kind, G):
    from oracle import oracle as O
    from mrpop import scenes
    from mrpop._lib import gpu_checker
    p = scenes.make_problem(kind, 11, gpu_checker(ctx), n_goals=7)
    rng = np.random.default_rng(G)
    starts = np.concatenate([p.start[None], p.goals])
    frm, to = _random_edges(rng, starts, 96, p.dim, 0.01, 2.0)
    print(f"[gpu] live {kind} G={G}", flush=True)
    w = ctx.world(p.scene, p.n_arms, p.bases)
    for mode in (O.MODE_FULL, O.MODE_DRRT):
        edge, mask = w.check_edges(mode, G, frm, to, with_mask=True)
        r = O.check_edges(p.scene, mode, p.bases, G, frm, to)
        _assert_masks(mask, edge, r["mask_lo"], r["mask_hi"], f"{kind} mode {mode} G {G}")
        for probe in (1, 4, 5):
            assert np.array_equal(_check_probe(w, mode, G, frm, to, probe), edge), (mode, probe)
    for arm in range(p.n_arms):
        f1, t1 = frm[:, 7 * arm:7 * arm + 7], to[:, 7 * arm:7 * arm + 7]
        edge, mask = w.check_edges(O.MODE_SINGLE, G, f1, t1, arm=arm, with_mask=True)
        r = O.check_edges(p.scene, O.MODE_SINGLE, p.bases, G, f1, t1, arm=arm)
        _assert_masks(mask, edge, r["mask_lo"], r["mask_hi"], f"{kind} single arm {arm} G {G}")


def _crowded_scene(rng, n):
    """n small obstacles of all six record types (spheres, capsules,
    z-capsules, world-axis cuboids, rotated cuboids, z-cuboids), in turn, in
    a box that grows with n."""
    from mrpop import scenes
    sph, cap, zcap, cub, zcub = [], [], [], [], []
    k = (n / 64.0) ** 0.5
    for i in range(n):
        c = rng.uniform((-1.2 * k, -1.2 * k, 0.0), (1.2 * k, 1.2 * k, 1.3 * k))
        s = rng.uniform(0.01, 0.03)
        kind = i % 6
        if kind == 0:
            sph.append(scenes.sphere_record(c, s))
        elif kind == 1:
            cap.append(scenes.capsule_record(c, tuple(rng.uniform(-np.pi, np.pi, 3)), s, rng.uniform(0.02, 0.15)))
        elif kind == 2:
            h = rng.uniform(0.02, 0.2)
            zcap.append([c[0], c[1], c[2], 0.0, 0.0, h, s, 1.0 / (h * h)])
        elif kind == 3:
            cub.append(scenes.cuboid_record(c, (0.0, 0.0, 0.0), rng.uniform(0.01, 0.04, 3)))
        elif kind == 4:
            cub.append(scenes.cuboid_record(c, tuple(rng.uniform(-np.pi, np.pi, 3)), rng.uniform(0.01, 0.04, 3)))
        else:
            yaw = rng.uniform(-np.pi, np.pi)
            cz, sz = np.cos(yaw), np.sin(yaw)
            hx, hy, hz = rng.uniform(0.01, 0.04, 3)
            zcub.append([c[0], c[1], c[2], cz, sz, 0.0, -sz, cz, 0.0, 0.0, 0.0, 1.0, hx, hy, hz])
    f = lambda rows, k: np.array(rows, np.float32).reshape(-1, k)
    return scenes.Scene(f(sph, 4), f(cap, 8), f(zcap, 8), f(cub, 15), f(zcub, 15))


@pytest.mark.parametrize("n_obstacles", [63, 64, 65, 150])
def test_crowded_scene_parity(ctx, n_obstacles):
    """Worlds around the checker's 64-obstacle chunk: per-chunk candidate
    masks up to 64 obstacles, several chunks and a full fine walk above."""
    from oracle import oracle as O
    from mrpop import scenes
    from mrpop._lib import gpu_checker
    p = scenes.make_problem("bin4", 11, gpu_checker(ctx), n_goals=7)
    rng = np.random.default_rng(n_obstacles)
    scene = _crowded_scene(rng, n_obstacles)
    assert scene.n_obstacles == n_obstacles
    starts = np.concatenate([p.start[None], p.goals])
    frm, to = _random_edges(rng, starts, 128, p.dim, 0.02, 0.8)
    print(f"[gpu] crowded scene n={n_obstacles}", flush=True)
    w = ctx.world(scene, p.n_arms, p.bases)
    G = 32
    masks = {}
    for mode in (O.MODE_FULL, O.MODE_DRRT):
        edge, mask = w.check_edges(mode, G, frm, to, with_mask=True)
        r = O.check_edges(scene, mode, p.bases, G, frm, to)
        _assert_masks(mask, edge, r["mask_lo"], r["mask_hi"], f"crowded {n_obstacles} mode {mode}")
        assert np.array_equal(w.check_edges(mode, G, frm, to), edge)
        masks[mode] = mask.astype(np.uint64)
    # samples colliding only with the environment, and free samples, both occur
    env_only = np.unpackbits((masks[O.MODE_FULL] & ~masks[O.MODE_DRRT]).view(np.uint8)).mean()
    free = 1.0 - np.unpackbits(masks[O.MODE_FULL].view(np.uint8)).mean() * 64 / G
    assert env_only > 0.05 / 2 and free > 0.05, (env_only, free)
    f1, t1 = frm[:, 7:14], to[:, 7:14]
    edge, mask = w.check_edges(O.MODE_SINGLE, G, f1, t1, arm=1, with_mask=True)
    r = O.check_edges(scene, O.MODE_SINGLE, p.bases, G, f1, t1, arm=1)
    _assert_masks(mask, edge, r["mask_lo"], r["mask_hi"], f"crowded {n_obstacles} single")


def test_fixture_configs_free_on_gpu(ctx):
    """Every start/goal of the reference's two problem files is collision-free
    (the files mark all 100 problems valid)."""
    from mrpop import scenes, geometry
    for fname, A in (("panda_four", 4), ("panda_five", 5)):
        probs = scenes.load_problem_file(GOLDEN / f"{fname}_problems.json", geometry.default_bases(A))
        w = ctx.world(probs[0].scene, A)
        cfgs = np.concatenate([np.stack([p.start for p in probs]), np.concatenate([p.goals for p in probs])])
        assert not w.config_collides(cfgs).any(), fname


@pytest.mark.parametrize("A", [2, 3, 4, 5])
def test_sphere_centres_vs_fp64(ctx, A):
    from oracle import oracle as O
    from mrpop import scenes, geometry
    bases = geometry.default_bases(A)
    rng = np.random.default_rng(A)
    u = rng.uniform(0, 1, (64, 7 * A)).astype(np.float32)
    q = (u * np.tile(scenes.SCALE, A) + np.tile(scenes.LOW, A)).astype(np.float32)
    w = ctx.world(scenes.Scene(), A, bases)
    co, fi = w.sphere_centres(q)
    worst = 0.0
    for i in range(len(q)):
        c64, f64 = O.sphere_centres(bases, q[i].astype(np.float64))
        worst = max(worst, np.abs(co[i] - c64).max(), np.abs(fi[i] - f64).max())
    assert worst < FK_TOL, worst


def test_empty_and_ragged_batches(ctx):
    from mrpop import scenes, geometry
    w = ctx.world(scenes.Scene(cuboids=[scenes.cuboid_record((0.5, 0, 0.5), (0, 0, 0), (0.1, 0.1, 0.1))]), 4,
                  geometry.default_bases(4))
    z = np.zeros((0, 28), np.float32)
    assert w.check_edges(0, 32, z, z).shape == (0,)
    rng = np.random.default_rng(0)
    # edge counts that leave partly filled waves for every samples-per-wave value
    for n in (1, 3, 13, 17, 65):
        frm = rng.uniform(-1, 1, (n, 28)).astype(np.float32)
        to = rng.uniform(-1, 1, (n, 28)).astype(np.float32)
        for G in (1, 5, 31, 64):
            e1, m1 = w.check_edges(0, G, frm, to, with_mask=True)
            # same edges one at a time
            for i in range(0, n, max(1, n // 4)):
                e2, m2 = w.check_edges(0, G, frm[i:i + 1], to[i:i + 1], with_mask=True)
                assert e2[0] == e1[i] and m2[0] == m1[i]


def test_invalid_arguments(ctx):
    from mrpop import scenes, _lib
    w = ctx.world(scenes.Scene(), 4)
    one = np.zeros((1, 28), np.float32)
    with pytest.raises(_lib.MrpopError):
        w.check_edges(0, 65, one, one)
    with pytest.raises(_lib.MrpopError):
        w.check_edges(0, 0, one, one)
    with pytest.raises(_lib.MrpopError):
        w.check_edges(_lib.MODE_SINGLE, 8, one[:, :7], one[:, :7], arm=4)

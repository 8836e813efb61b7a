"""CPU tests pinning the FP64 oracle (oracle/oracle.cpp) before it is trusted
as the GPU checker's reference.

What pins what:
  * geometry tables      == the numbers in the reference's robot headers
                           (tests/golden/panda_tables.json, extracted from text);
  * forward kinematics   == textbook Franka Panda kinematics (modified DH,
                           Craig convention) -- independent of the tables;
  * primitive clearances == brute force over densely sampled obstacles;
  * composite verdict    == an independent numpy restatement written as
                           explicit pair lists from the reference text, incl.
                           the multi-arm intra-arm quirk (panda_four.cuh:390-401),
                           which the test shows is not a no-op;
  * known answers        == all 200 start/goal configurations of the reference's
                           problem files are collision-free (all marked valid);
  * invariants           rigid motion of the whole world leaves verdicts
                           unchanged; inflating obstacles never frees a sample.
"""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import oracle as O
from mrpop import scenes, geometry

TABLES = json.loads((GOLDEN / "panda_tables.json").read_text())


def test_tables_match_reference_text():
    t = O.tables()
    ft = np.array(TABLES["fine_fixed_transforms"], np.float32).reshape(8, 4, 4)
    assert np.array_equal(t["joint_tf"], ft[:, :3, :].reshape(8, 12))
    assert np.array_equal(t["coarse"], np.array(TABLES["coarse_spheres"], np.float32).reshape(11, 4))
    assert np.array_equal(t["fine"], np.array(TABLES["fine_spheres"], np.float32).reshape(59, 4))
    assert t["coarse_joint"].tolist() == [int(x) for x in TABLES["coarse_sphere_to_joint"]]
    assert t["fine_joint"].tolist() == [int(x) for x in TABLES["fine_sphere_to_joint"]]
    assert t["coarse_ranges"].reshape(-1).tolist() == [int(x) for x in TABLES["coarse_self_cc_ranges"]]
    assert t["fine_ranges"].reshape(-1).tolist() == [int(x) for x in TABLES["fine_self_cc_ranges"]]
    assert np.array_equal(t["scale"], np.array(TABLES["scale_s_m"], np.float32))
    assert np.array_equal(t["low"], np.array(TABLES["scale_s_a"], np.float32))
    for A in (4, 5):
        ref = np.array(TABLES[f"base_{A}"], np.float32).reshape(A, 4, 4)
        assert np.array_equal(t[f"base{A}"], ref[:, :3, :].reshape(A, 12))
        assert np.array_equal(geometry.default_bases(A), ref[:, :3, :].reshape(A, 12))
    # joint types: every moving joint is a z rotation; joint 0 is the base
    assert [int(x) for x in TABLES["fine_joint_types"]][1:] == [5] * 7
    assert [int(x) for x in TABLES["fine_joint_id_to_dof"]] == [-1, 0, 1, 2, 3, 4, 5, 6]


def _dh_frames(q):
    """Franka Panda, modified DH (Craig): T_i = Rx(alpha) Tx(a) Rz(q_i) Tz(d)."""
    a = [0, 0, 0, 0.0825, -0.0825, 0, 0.088]
    d = [0.333, 0, 0.316, 0, 0.384, 0, 0]
    alpha = [0, -math.pi / 2, math.pi / 2, math.pi / 2, -math.pi / 2, math.pi / 2, math.pi / 2]
    T = np.eye(4)
    out = [T.copy()]
    for i in range(7):
        ca, sa = math.cos(alpha[i]), math.sin(alpha[i])
        cq, sq = math.cos(q[i]), math.sin(q[i])
        rx = np.array([[1, 0, 0, 0], [0, ca, -sa, 0], [0, sa, ca, 0], [0, 0, 0, 1]])
        tx = np.eye(4)
        tx[0, 3] = a[i]
        rz = np.array([[cq, -sq, 0, 0], [sq, cq, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
        tz = np.eye(4)
        tz[2, 3] = d[i]
        T = T @ rx @ tx @ rz @ tz
        out.append(T.copy())
    return out


def test_fk_matches_textbook_panda_kinematics():
    ident = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    f0 = O.arm_frames(ident, np.zeros(7))
    assert np.allclose(f0[7][:, 3], [0.088, 0.0, 1.033], atol=1e-12)
    rng = np.random.default_rng(0)
    for _ in range(50):
        q = rng.uniform(-3, 3, 7)
        fr = O.arm_frames(ident, q)
        dh = _dh_frames(q)
        for j in range(8):
            assert np.allclose(fr[j], dh[j][:3, :], atol=1e-6), j   # tables hold 4-5 significant digits
            R = fr[j][:, :3]
            assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)


def test_fk_base_pose_and_rigid_links():
    rng = np.random.default_rng(1)
    bases = geometry.default_bases(4)
    q = rng.uniform(-2, 2, 28)
    co, fi = O.sphere_centres(bases, q)
    ident = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    for a in range(4):
        _, fi1 = O.sphere_centres(ident, q[7 * a:7 * a + 7])
        B = bases[a].reshape(3, 4).astype(np.float64)
        assert np.allclose(fi1[0] @ B[:, :3].T + B[:, 3], fi[a], atol=1e-12)
    # spheres on one link keep their distances for any q
    _, fa = O.sphere_centres(ident, rng.uniform(-2, 2, 7))
    _, fb = O.sphere_centres(ident, rng.uniform(-2, 2, 7))
    jt = np.array(TABLES["fine_sphere_to_joint"]).astype(int)
    for j in range(8):
        idx = np.where(jt == j)[0]
        da = np.linalg.norm(fa[0][idx, None] - fa[0][None, idx], axis=-1)
        db = np.linalg.norm(fb[0][idx, None] - fb[0][None, idx], axis=-1)
        assert np.allclose(da, db, atol=1e-12)


def test_reference_fixture_configs_are_free():
    worst = np.inf
    for fname, A in (("panda_four", 4), ("panda_five", 5)):
        probs = scenes.load_problem_file(GOLDEN / f"{fname}_problems.json", geometry.default_bases(A))
        assert len(probs) == 50
        for p in probs:
            for q in [p.start] + list(p.goals):
                nom, lo, hi, m = O.check_config(p.scene, O.MODE_FULL, p.bases, q)
                assert not nom and not hi, p.name
                worst = min(worst, m)
    assert worst > O.EPS_CONTACT


def test_problem_file_bases_match_compiled_tables():
    for fname, A in (("panda_four", 4), ("panda_five", 5)):
        file_bases = scenes.load_problem_file(GOLDEN / f"{fname}_problems.json")[0].bases
        assert np.abs(file_bases - geometry.default_bases(A)).max() < 1e-6


# ------------------------------------------------------------- primitives --
def _brute_clearance(points, p, r):
    return np.min(np.linalg.norm(points - p, axis=1)) - r


def _box_points(rec, n=40):
    c, axes, h = rec[0:3], np.stack([rec[3:6], rec[6:9], rec[9:12]]), rec[12:15]
    g = np.linspace(-1, 1, n)
    u = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * h
    return c + u @ axes, 2 * h.max() / (n - 1)


def _capsule_points(rec, n=400, m=24):
    p1, v, r = rec[0:3].astype(np.float64), rec[3:6].astype(np.float64), float(rec[6])
    t = np.linspace(0, 1, n)[:, None]
    axis = p1 + t * v
    # the capsule's distance field = segment distance - radius
    return axis, np.linalg.norm(v) / (n - 1), r


def test_primitive_clearances_vs_brute_force():
    rng = np.random.default_rng(2)
    for _ in range(60):
        box = scenes.cuboid_record(rng.uniform(-0.3, 0.3, 3), rng.uniform(-math.pi, math.pi, 3),
                                   rng.uniform(0.02, 0.2, 3))
        pts, step = _box_points(box)
        cap = scenes.capsule_record(rng.uniform(-0.3, 0.3, 3), rng.uniform(-math.pi, math.pi, 3),
                                    rng.uniform(0.01, 0.1), rng.uniform(0.05, 0.4))
        axis, astep, crad = _capsule_points(cap)
        sph = scenes.sphere_record(rng.uniform(-0.3, 0.3, 3), rng.uniform(0.01, 0.2))
        for _ in range(8):
            p = rng.uniform(-0.6, 0.6, 3)
            r = rng.uniform(0.01, 0.15)
            got = O.env_clearance(scenes.Scene(cuboids=[box]), p, r)
            want = _brute_clearance(pts, p, r)
            inside = np.all(np.abs((p - box[0:3]) @ np.stack([box[3:6], box[6:9], box[9:12]]).T) <= box[12:15])
            if inside:
                assert math.isclose(got, -r, abs_tol=1e-9)
            else:
                assert want - step * 0.87 - 1e-6 <= got <= want + 1e-6
            got = O.env_clearance(scenes.Scene(capsules=[cap]), p, r)
            want = _brute_clearance(axis, p, r + crad)
            assert want - astep - 1e-6 <= got <= want + 1e-6
            got = O.env_clearance(scenes.Scene(spheres=[sph]), p, r)
            assert math.isclose(got, np.linalg.norm(p - sph[:3]) - r - sph[3], abs_tol=1e-7)


def test_zero_length_capsule_is_a_sphere_and_z_variants():
    rng = np.random.default_rng(3)
    for _ in range(50):
        c = rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        # a vertical capsule is tested identically by the general and z-aligned formulas
        capz = scenes.capsule_record_endpoints(c, c + np.array([0, 0, rng.uniform(0.05, 0.3)], np.float32), 0.04)
        p = rng.uniform(-0.7, 0.7, 3)
        g = O.env_clearance(scenes.Scene(capsules=[capz]), p, 0.05)
        z = O.env_clearance(scenes.Scene(z_capsules=[capz]), p, 0.05)
        assert math.isclose(g, z, abs_tol=1e-7)
        box = scenes.cuboid_record(c, (0, 0, rng.uniform(-3, 3)), rng.uniform(0.02, 0.2, 3))
        g = O.env_clearance(scenes.Scene(cuboids=[box]), p, 0.05)
        z = O.env_clearance(scenes.Scene(z_cuboids=[box]), p, 0.05)
        assert math.isclose(g, z, abs_tol=1e-6)


# ----------------------------------------------- independent restatement --
def _env_clear_np(scene, P, R):
    """Minimum clearance of spheres (P[n,3], R[n]) against the scene (numpy)."""
    m = np.full(len(P), np.inf)
    for s in scene.spheres:
        m = np.minimum(m, np.linalg.norm(P - s[:3], axis=1) - R - s[3])
    for c in scene.capsules:
        p1, v = c[0:3].astype(np.float64), c[3:6].astype(np.float64)
        t = np.clip(((P - p1) @ v) * c[7], 0, 1)
        m = np.minimum(m, np.linalg.norm(P - (p1 + t[:, None] * v), axis=1) - R - c[6])
    for b in scene.cuboids:
        ax = np.stack([b[3:6], b[6:9], b[9:12]]).astype(np.float64)
        d = np.maximum(np.abs((P - b[0:3]) @ ax.T) - b[12:15], 0)
        m = np.minimum(m, np.linalg.norm(d, axis=1) - R)
    return m


def _composite_np(scene, bases, q, mode, quirk=True):
    """Verdict written as explicit pair lists from the reference text."""
    cj = np.array(TABLES["coarse_sphere_to_joint"]).astype(int)
    fj = np.array(TABLES["fine_sphere_to_joint"]).astype(int)
    cr = np.array(TABLES["coarse_spheres"]).reshape(11, 4)[:, 3]
    fr = np.array(TABLES["fine_spheres"]).reshape(59, 4)[:, 3]
    crange = np.array(TABLES["coarse_self_cc_ranges"]).astype(int).reshape(5, 3)
    frange = np.array(TABLES["fine_self_cc_ranges"]).astype(int).reshape(25, 3)
    A = len(bases)
    co, fi = O.sphere_centres(bases, q)
    env = np.zeros((A, 8), bool)
    slf = np.zeros((A, 8), bool)
    if mode != O.MODE_DRRT:
        for a in range(A):
            hit = _env_clear_np(scene, co[a][1:], cr[1:]) < 0
            env[a, cj[1:][hit]] = True
    intra_c = [(s1, j) for s1, lo, hi in crange for j in range(lo, hi + 1)]
    for a in range(A):
        for s1, j in intra_c:
            if np.linalg.norm(co[a][s1] - co[a][j]) < cr[s1] + cr[j]:
                slf[a, cj[s1]] = True
        for b in range(a + 1, A):
            d = np.linalg.norm(co[a][:, None] - co[b][None], axis=-1) < cr[:, None] + cr[None]
            for i in np.where(d.any(axis=1))[0]:
                slf[a, cj[i]] = True
    # fine
    if mode != O.MODE_DRRT:
        for a in range(A):
            idx = [i for i in range(1, 59) if env[a, fj[i]]]
            if idx and (_env_clear_np(scene, fi[a][idx], fr[idx]) < 0).any():
                return True
    if mode == O.MODE_SINGLE or not quirk:
        intra_f = [(s1, s1, fj[s1], j) for s1, lo, hi in frange for j in range(lo, hi + 1)]
    else:  # multi-arm: range r gated by / centred on fine sphere r, radius of ranges[r][0]
        intra_f = [(r, frange[r][0], fj[r], j) for r in range(25) for j in range(frange[r][1], frange[r][2] + 1)]
    for a in range(A):
        for pos, rad, gate, j in intra_f:
            if slf[a, gate] and np.linalg.norm(fi[a][pos] - fi[a][j]) < fr[rad] + fr[j]:
                return True
        for b in range(a + 1, A):
            for i in range(59):
                if slf[a, fj[i]] and (np.linalg.norm(fi[a][i] - fi[b], axis=1) < fr[i] + fr).any():
                    return True
    return False


def _random_configs(rng, n, A, near=None, spread=0.6):
    if near is None:
        u = rng.uniform(0, 1, (n, 7 * A))
        return u * np.tile(scenes.SCALE, A) + np.tile(scenes.LOW, A)
    idx = rng.integers(len(near), size=n)
    return near[idx] + rng.normal(scale=spread, size=(n, near.shape[1]))


def test_composite_matches_independent_restatement():
    rng = np.random.default_rng(4)
    n_checked = n_hit = 0
    for fname, A in (("panda_four", 4), ("panda_five", 5)):
        probs = scenes.load_problem_file(GOLDEN / f"{fname}_problems.json", geometry.default_bases(A))
        starts = np.stack([p.start for p in probs])
        for q in _random_configs(rng, 60, A, near=starts, spread=0.5):
            for mode in (O.MODE_FULL, O.MODE_DRRT):
                nom, lo, hi, _ = O.check_config(probs[0].scene, mode, probs[0].bases, q)
                if lo != hi:
                    continue
                assert nom == _composite_np(probs[0].scene, probs[0].bases, q, mode), (fname, mode)
                n_checked += 1
                n_hit += nom
    assert n_checked > 150 and 0.1 < n_hit / n_checked < 0.95


def test_multi_arm_quirk_is_reproduced_and_matters():
    """The multi-arm fine self check uses fine sphere r (not ranges[r][0]) for
    ranges 6..24.  Find configurations where that changes the verdict and check
    the oracle gives the reference's (quirk) answer there."""
    rng = np.random.default_rng(5)
    bases = geometry.default_bases(4)
    empty = scenes.Scene()
    differ = 0
    for _ in range(400):
        q = _random_configs(rng, 1, 4)[0]
        nom, lo, hi, _ = O.check_config(empty, O.MODE_DRRT, bases, q)
        if lo != hi:
            continue
        quirk = _composite_np(empty, bases, q, O.MODE_DRRT, quirk=True)
        plain = _composite_np(empty, bases, q, O.MODE_DRRT, quirk=False)
        assert nom == quirk
        differ += quirk != plain
        if differ >= 3:
            break
    assert differ >= 3, "no configuration distinguishes the quirk: the test would not catch its loss"


def test_single_arm_uses_correct_ranges():
    rng = np.random.default_rng(6)
    bases = geometry.default_bases(4)
    probs = scenes.load_problem_file(GOLDEN / "panda_four_problems.json", bases)
    for _ in range(150):
        q = _random_configs(rng, 1, 1)[0]
        nom, lo, hi, _ = O.check_config(probs[0].scene, O.MODE_SINGLE, bases, q, arm=2)
        if lo == hi:
            assert nom == _composite_np(probs[0].scene, bases[2:3], q, O.MODE_SINGLE)


def test_rigid_motion_invariance():
    rng = np.random.default_rng(7)
    p = scenes.make_problem("bin4", 3, O.config_checker, n_goals=3)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    t = np.array([0.3, -0.2, 0.1])
    moved_scene = p.scene.transformed(R, t)
    B = p.bases.reshape(-1, 3, 4).astype(np.float64)
    moved_bases = np.concatenate([np.einsum("ij,ajk->aik", R, B[:, :, :3]),
                                  (np.einsum("ij,aj->ai", R, B[:, :, 3]) + t)[:, :, None]], axis=2).reshape(-1, 12)
    starts = np.concatenate([p.start[None], p.goals])
    frm = starts[rng.integers(len(starts), size=64)]
    to = (frm + rng.normal(scale=0.3, size=frm.shape)).astype(np.float32)
    a = O.check_edges(p.scene, O.MODE_FULL, p.bases, 16, frm, to)
    b = O.check_edges(moved_scene, O.MODE_FULL, moved_bases, 16, frm, to, eps=1e-4)
    robust = (a["mask_lo"] == a["mask_hi"]) & (b["mask_lo"] == b["mask_hi"])
    assert robust.mean() > 0.9
    assert np.array_equal(a["mask_nom"][robust], b["mask_nom"][robust])


def test_inflation_is_monotone():
    rng = np.random.default_rng(8)
    p = scenes.make_problem("shelf5", 2, O.config_checker, n_goals=3)
    starts = np.concatenate([p.start[None], p.goals])
    frm = starts[rng.integers(len(starts), size=48)]
    to = (frm + rng.normal(scale=0.3, size=frm.shape)).astype(np.float32)
    a = O.check_edges(p.scene, O.MODE_FULL, p.bases, 16, frm, to)
    b = O.check_edges(p.scene.inflated(0.01), O.MODE_FULL, p.bases, 16, frm, to)
    assert np.all((a["mask_nom"] & ~b["mask_nom"]) == 0)
    assert (b["mask_nom"] != a["mask_nom"]).any()


def test_fp32_mirror_agrees_on_robust_edges():
    for f in sorted(GOLDEN.glob("edges_*.npz")):
        z = np.load(f)
        scene = scenes.Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])
        got = O.check_edges_f32(scene, int(z["mode"]), z["bases"], int(z["G"]), z["frm"], z["to"], arm=int(z["arm"]),
                                threads=4)
        robust = (z["mask_lo"] != 0) == (z["mask_hi"] != 0)
        assert np.array_equal(got[robust], (z["mask_lo"] != 0)[robust]), f.name


def test_golden_vectors_reproduce():
    for f in sorted(GOLDEN.glob("edges_*.npz")):
        z = np.load(f)
        scene = scenes.Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])
        r = O.check_edges(scene, int(z["mode"]), z["bases"], int(z["G"]), z["frm"], z["to"], arm=int(z["arm"]))
        for k in ("mask_nom", "mask_lo", "mask_hi"):
            assert np.array_equal(r[k], z[k]), (f.name, k)


def test_quirk_golden_distinguishes_the_reading():
    """edges_panda_four_quirk_drrt_g16.npz: every edge ends at a robust
    configuration whose verdict the reference's multi-arm fine-range reading
    (panda_four.cuh:390-401) flips against the single-arm reading, and that
    reading is reproduced; so a device that read the ranges the single-arm
    way fails the golden comparison on every edge's last sample."""
    z = np.load(GOLDEN / "edges_panda_four_quirk_drrt_g16.npz")
    scene = scenes.Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])
    last = np.uint64(1) << np.uint64(int(z["G"]) - 1)
    assert ((z["mask_nom"] ^ z["mask_plain"]) & last != 0).all()
    assert ((z["mask_lo"] ^ z["mask_hi"]) == 0).all()
    r = O.check_edges_plain_ranges(scene, int(z["mode"]), z["bases"], int(z["G"]), z["frm"], z["to"])
    assert np.array_equal(r["mask_nom"], z["mask_plain"])


def test_zaligned_golden_holds_the_z_obstacles():
    """The z-aligned golden scene holds z-capsules and z-cuboids that the
    edges actually hit (a device ignoring them would miss collisions)."""
    for name in ("edges_posts3_full_g32.npz", "edges_posts3_single_g64.npz"):
        z = np.load(GOLDEN / name)
        assert len(z["z_capsules"]) and len(z["z_cuboids"])
        full = scenes.Scene(z["spheres"], z["capsules"], z["z_capsules"], z["cuboids"], z["z_cuboids"])
        bare = scenes.Scene(z["spheres"], z["capsules"], np.zeros((0, 8), np.float32), z["cuboids"],
                            np.zeros((0, 15), np.float32))
        r = O.check_edges(bare, int(z["mode"]), z["bases"], int(z["G"]), z["frm"], z["to"], arm=int(z["arm"]))
        assert ((z["mask_nom"] & ~r["mask_nom"]) != 0).sum() >= 20, name


def test_edge_interpolation_excludes_start_includes_end():
    """Sample k = 1..G of an edge: a free start with a colliding end must
    collide at sample G only when the midway samples are free."""
    probs = scenes.load_problem_file(GOLDEN / "panda_four_problems.json", geometry.default_bases(4))
    p = probs[0]
    box = scenes.cuboid_record((0, 0, 2.0), (0, 0, 0), (0.01, 0.01, 0.01))   # far away: a free world
    scene = scenes.Scene(cuboids=np.concatenate([p.scene.cuboids, box[None]]))
    frm = p.start[None].astype(np.float32)
    r = O.check_edges(scene, O.MODE_FULL, p.bases, 8, frm, frm)
    assert r["mask_nom"][0] == 0
    # a zero-length edge samples its end point G times
    bad = frm.copy()
    bad[0, 1] = 1.5   # swing arm 0's shoulder down into the table
    r = O.check_edges(scene, O.MODE_FULL, p.bases, 4, bad, bad)
    assert r["mask_nom"][0] in (0, 0b1111)

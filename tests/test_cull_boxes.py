"""The checker skips the test of a coarse sphere against an obstacle when
the box of the wave's centres of that sphere misses the obstacle's cull box
grown by r * f (r the sphere's radius, f stored with the box).  That is exact
only if every centre a test reports as colliding lies in the grown box.
Checked here on the CPU against FP64 restatements of the reference's tests
(utils.cuh:100-229) with random obstacles of every type, rotated cuboids
included, points sampled around and inside them, and radii up to 0.2 m (the
largest coarse sphere is 0.176 m)."""
import ctypes

import numpy as np
import pytest

from mrpop import _lib, scenes


def _tests(kind, o, p, r):
    """FP64 restatement of the test value (< 0: collision) for points p (n, 3)."""
    p = p.astype(np.float64)
    o = o.astype(np.float64)
    if kind == "sphere":
        d = p - o[:3]
        return (d * d).sum(1) - (r + o[3]) ** 2
    if kind in ("capsule", "zcapsule"):
        v = o[3:6] if kind == "capsule" else np.array([0, 0, o[5]])
        if kind == "capsule":
            dot = (p - o[:3]) @ v
        else:
            dot = (p[:, 2] - o[2]) * o[5]
        t = np.clip(dot * o[7], 0, 1)
        c = o[:3] + t[:, None] * v
        d = p - c
        return (d * d).sum(1) - (r + o[6]) ** 2
    if kind in ("cuboid", "zcuboid"):
        A = o[3:12].reshape(3, 3).copy()
        if kind == "zcuboid":
            A[0, 2] = A[1, 2] = 0
            A[2] = [0, 0, 1]
        proj = np.abs((p - o[:3]) @ A.T) - o[12:15]
        a = np.maximum(proj, 0)
        return (a * a).sum(1) - r * r
    raise ValueError(kind)


def _random_scene(rng):
    sph = [scenes.sphere_record(rng.uniform(-1, 1, 3), rng.uniform(0.01, 0.2)) for _ in range(6)]
    caps = []
    for _ in range(6):
        p, q = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        caps.append(scenes.capsule_record_endpoints(p, q, rng.uniform(0.01, 0.1)))
    zcaps = []
    for _ in range(4):
        rec = np.zeros(8, np.float32)
        rec[:3] = rng.uniform(-1, 1, 3)
        rec[5] = rng.uniform(0.05, 0.5)
        rec[6] = rng.uniform(0.01, 0.1)
        rec[7] = 1.0 / (rec[5] * rec[5])
        zcaps.append(rec)
    cubs = [scenes.cuboid_record(rng.uniform(-1, 1, 3), rng.uniform(-np.pi, np.pi, 3) * (k % 2),
                                 rng.uniform(0.01, 0.3, 3)) for k in range(8)]
    zcubs = []
    for _ in range(4):
        a = rng.uniform(-np.pi, np.pi)
        rec = np.zeros(15, np.float32)
        rec[:3] = rng.uniform(-1, 1, 3)
        rec[3:12] = [np.cos(a), np.sin(a), 0, -np.sin(a), np.cos(a), 0, 0, 0, 1]
        rec[12:15] = rng.uniform(0.01, 0.3, 3)
        zcubs.append(rec)
    return scenes.Scene(spheres=np.array(sph, np.float32).reshape(-1, 4),
                        capsules=np.array(caps, np.float32).reshape(-1, 8),
                        z_capsules=np.array(zcaps, np.float32).reshape(-1, 8),
                        cuboids=np.array(cubs, np.float32).reshape(-1, 15),
                        z_cuboids=np.array(zcubs, np.float32).reshape(-1, 15))


def _checker_order(scene):
    """Obstacles in the checker's order: world-axis cuboids before the others."""
    out = [("sphere", o) for o in scene.spheres] + [("capsule", o) for o in scene.capsules]
    out += [("zcapsule", o) for o in scene.z_capsules]
    axis = [o for o in scene.cuboids if np.array_equal(o[3:12], np.eye(3, dtype=np.float32).ravel())]
    other = [o for o in scene.cuboids if not np.array_equal(o[3:12], np.eye(3, dtype=np.float32).ravel())]
    out += [("cuboid", o) for o in axis] + [("cuboid", o) for o in other] + [("zcuboid", o) for o in scene.z_cuboids]
    return out


@pytest.mark.parametrize("seed", range(4))
def test_cull_boxes_hold_every_colliding_centre(seed):
    rng = np.random.default_rng(seed)
    scene = _random_scene(rng)
    lib = _lib.lib()
    lib.mrpop_diag_world_boxes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    cs = scene.as_c()
    n = lib.mrpop_diag_world_boxes(ctypes.byref(cs), None, 0)
    obs = _checker_order(scene)
    assert n == len(obs)
    boxes = np.zeros((n, 8), np.float32)
    lib.mrpop_diag_world_boxes(ctypes.byref(cs), boxes.ctypes.data, n)
    for (kind, o), box in zip(obs, boxes):
        lo, hi, f = box[:3].astype(np.float64), box[3:6].astype(np.float64), float(box[6])
        assert 1.0 <= f < 1.8  # rotations: sum of |entries| of a row is at most sqrt(3)
        centre = (lo + hi) / 2
        half = (hi - lo) / 2 + 0.2 * f
        # points around the grown box and well beyond it
        p = centre + rng.uniform(-1.6, 1.6, (20000, 3)) * np.maximum(half, 0.05)
        r = rng.uniform(0.0, 0.2, (20000,))
        hit = _tests(kind, o, p, r) < 0
        g = (r * f)[:, None]
        inside = ((p >= lo - g) & (p <= hi + g)).all(1)
        assert hit.any(), kind
        assert not (hit & ~inside).any(), f"{kind}: colliding centre outside its cull box"
        assert (~inside).any()

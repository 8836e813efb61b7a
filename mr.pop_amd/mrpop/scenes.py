"""Scenes and planning problems: the reference's problem-file schema and the
synthetic scene generators used for benchmarking.

Scene construction follows the reference's factory and harness:
  * Euler XYZ orientation -> R = Rz(phi) * Ry(theta) * Rx(rho) and cuboid axes
    = columns of R (src/collision/factory.hh:26-41);
  * cylinder/capsule from centre + Euler + length: endpoints at +-length/2 on
    the local z axis, rdv = 1 / |v|^2 (factory.hh:100-114, 144-153);
  * problem-file mapping: "sphere" -> sphere, "cylinder" -> capsule (the
    reference checks cylinders as capsules), "box" -> cuboid
    (scripts/evaluate_mr.cu:14-79).  The file's base poses are not used by the
    reference (its base transforms are compiled in); here they are kept and
    compared against the compiled-in defaults by the tests.

Records are float32, laid out as the C ABI's mrpop_scene expects
(include/mrpop_c.h).
"""
from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, List, Optional, Sequence

import numpy as np

DOF_PER_ARM = 7
# Joint limits (src/planning/Robots.hh:13-23): q = u * SCALE + LOW.
SCALE = np.array([5.9342, 3.6652, 5.9342, 3.2289, 5.9342, 3.9095999999999997, 5.9342], np.float32)
LOW = np.array([-2.9671, -1.8326, -2.9671, -3.1416, -2.9671, -0.0873, -2.9671], np.float32)


def euler_xyz_matrix(rho: float, theta: float, phi: float) -> np.ndarray:
    """R = Rz(phi) * Ry(theta) * Rx(rho), in float32 (factory.hh:29-31)."""
    cr, sr = np.float32(math.cos(rho)), np.float32(math.sin(rho))
    ct, st = np.float32(math.cos(theta)), np.float32(math.sin(theta))
    cp, sp = np.float32(math.cos(phi)), np.float32(math.sin(phi))
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]], np.float32)
    ry = np.array([[ct, 0, st], [0, 1, 0], [-st, 0, ct]], np.float32)
    rz = np.array([[cp, -sp, 0], [sp, cp, 0], [0, 0, 1]], np.float32)
    return (rz @ ry @ rx).astype(np.float32)


def cuboid_record(center, euler_xyz, half_extents) -> np.ndarray:
    r = euler_xyz_matrix(*euler_xyz)
    rec = np.zeros(15, np.float32)
    rec[0:3] = center
    rec[3:6] = r[:, 0]
    rec[6:9] = r[:, 1]
    rec[9:12] = r[:, 2]
    rec[12:15] = half_extents
    return rec


def capsule_record_endpoints(p1, p2, radius) -> np.ndarray:
    p1 = np.asarray(p1, np.float32)
    p2 = np.asarray(p2, np.float32)
    v = (p2 - p1).astype(np.float32)
    dot = np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    rec = np.zeros(8, np.float32)
    rec[0:3] = p1
    rec[3:6] = v
    rec[6] = radius
    rec[7] = np.float32(1.0 / float(dot))
    return rec


def capsule_record(center, euler_xyz, radius, length) -> np.ndarray:
    r = euler_xyz_matrix(*euler_xyz)
    c = np.asarray(center, np.float32)
    half = np.float32(length) / np.float32(2)
    p1 = r @ np.array([0, 0, half], np.float32) + c
    p2 = r @ np.array([0, 0, -half], np.float32) + c
    return capsule_record_endpoints(p1, p2, radius)


def sphere_record(center, radius) -> np.ndarray:
    return np.array([center[0], center[1], center[2], radius], np.float32)


class _CScene(ctypes.Structure):
    _fields_ = [("n_spheres", ctypes.c_int), ("n_capsules", ctypes.c_int), ("n_z_capsules", ctypes.c_int),
                ("n_cuboids", ctypes.c_int), ("n_z_cuboids", ctypes.c_int),
                ("spheres", ctypes.c_void_p), ("capsules", ctypes.c_void_p), ("z_capsules", ctypes.c_void_p),
                ("cuboids", ctypes.c_void_p), ("z_cuboids", ctypes.c_void_p)]


@dataclass
class Scene:
    """Obstacle lists as float32 record arrays (see include/mrpop_c.h)."""
    spheres: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    capsules: np.ndarray = field(default_factory=lambda: np.zeros((0, 8), np.float32))
    z_capsules: np.ndarray = field(default_factory=lambda: np.zeros((0, 8), np.float32))
    cuboids: np.ndarray = field(default_factory=lambda: np.zeros((0, 15), np.float32))
    z_cuboids: np.ndarray = field(default_factory=lambda: np.zeros((0, 15), np.float32))

    def __post_init__(self):
        for name, w in (("spheres", 4), ("capsules", 8), ("z_capsules", 8), ("cuboids", 15), ("z_cuboids", 15)):
            a = np.ascontiguousarray(np.asarray(getattr(self, name), np.float32).reshape(-1, w))
            setattr(self, name, a)

    @property
    def n_obstacles(self) -> int:
        return sum(len(a) for a in (self.spheres, self.capsules, self.z_capsules, self.cuboids, self.z_cuboids))

    def as_c(self) -> _CScene:
        """ctypes mrpop_scene viewing this scene's arrays (keep the Scene alive)."""
        def p(a):
            return a.ctypes.data if len(a) else None
        return _CScene(len(self.spheres), len(self.capsules), len(self.z_capsules), len(self.cuboids),
                       len(self.z_cuboids), p(self.spheres), p(self.capsules), p(self.z_capsules),
                       p(self.cuboids), p(self.z_cuboids))

    def inflated(self, margin: float) -> "Scene":
        """Every obstacle grown by `margin` metres (radius or half-extents)."""
        s = Scene(self.spheres.copy(), self.capsules.copy(), self.z_capsules.copy(), self.cuboids.copy(),
                  self.z_cuboids.copy())
        s.spheres[:, 3] += margin
        s.capsules[:, 6] += margin
        s.z_capsules[:, 6] += margin
        s.cuboids[:, 12:15] += margin
        s.z_cuboids[:, 12:15] += margin
        return s

    def transformed(self, rot: np.ndarray, trans: np.ndarray) -> "Scene":
        """The scene moved by the rigid motion x -> rot @ x + trans (float32)."""
        rot = np.asarray(rot, np.float32)
        trans = np.asarray(trans, np.float32)
        sp = self.spheres.copy()
        sp[:, 0:3] = sp[:, 0:3] @ rot.T + trans
        cap = self.capsules.copy()
        cap[:, 0:3] = cap[:, 0:3] @ rot.T + trans
        cap[:, 3:6] = cap[:, 3:6] @ rot.T
        cub = self.cuboids.copy()
        cub[:, 0:3] = cub[:, 0:3] @ rot.T + trans
        for k in range(3):
            cub[:, 3 + 3 * k:6 + 3 * k] = cub[:, 3 + 3 * k:6 + 3 * k] @ rot.T
        if len(self.z_capsules) or len(self.z_cuboids):
            raise ValueError("z-aligned obstacles are not closed under general rotations")
        return Scene(sp, cap, self.z_capsules, cub, self.z_cuboids)


@dataclass
class Problem:
    """One planning query: arm layout, scene, start and goal configurations."""
    name: str
    n_arms: int
    bases: np.ndarray            # (n_arms, 12) world <- base, row-major 3x4
    scene: Scene
    start: np.ndarray            # (7 * n_arms,)
    goals: np.ndarray            # (n_goals, 7 * n_arms)
    seed: int = 0

    @property
    def dim(self) -> int:
        return DOF_PER_ARM * self.n_arms


# ------------------------------------------------------------------ bases --
def yaw_pose(x, y, z, yaw) -> np.ndarray:
    c, s = math.cos(yaw), math.sin(yaw)
    return np.array([c, -s, 0, x, s, c, 0, y, 0, 0, 1, z], np.float32)


def quat_wxyz_pose(pos, q) -> np.ndarray:
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.zeros(12, np.float32)
    m[[0, 1, 2, 4, 5, 6, 8, 9, 10]] = r.reshape(-1)
    m[[3, 7, 11]] = pos
    return m


# ------------------------------------------------------- problem files ----
def scene_from_problem_dict(problem: dict) -> Scene:
    """scripts/evaluate_mr.cu:14-79: spheres, cylinders as capsules, boxes as cuboids."""
    spheres = [sphere_record(o["position"], o["radius"]) for o in problem.get("sphere", [])]
    if problem.get("name") == "box":
        # evaluate_mr.cu:28-37 -- only for a problem literally named "box"
        cub_cyl = [cuboid_record(o["position"], o["orientation_euler_xyz"],
                                 [o["radius"], o["radius"], o["radius"] / 2.0]) for o in problem.get("cylinder", [])]
        capsules = []
    else:
        cub_cyl = []
        capsules = [capsule_record(o["position"], o["orientation_euler_xyz"], o["radius"], o["length"])
                    for o in problem.get("cylinder", [])]
    cuboids = cub_cyl + [cuboid_record(o["position"], o["orientation_euler_xyz"], o["half_extents"])
                         for o in problem.get("box", [])]
    return Scene(np.array(spheres, np.float32).reshape(-1, 4), np.array(capsules, np.float32).reshape(-1, 8),
                 np.zeros((0, 8), np.float32), np.array(cuboids, np.float32).reshape(-1, 15))


def load_problem_file(path, default_bases: Optional[np.ndarray] = None) -> List[Problem]:
    """Read a reference problem file (schema of scripts/panda_*_problems.json).

    Only problems with "valid": true are returned, as the reference's driver
    skips the others (evaluate_mr.cu:121-124).  Bases: `default_bases` when
    given (the reference ignores the file's poses), else the file's poses.
    """
    data = json.loads(Path(path).read_text())
    out = []
    for p in data:
        if not p.get("valid", True):
            continue
        n_arms = len(p["start"]) // DOF_PER_ARM
        if default_bases is not None:
            bases = np.asarray(default_bases, np.float32).reshape(n_arms, 12)
        else:
            bases = np.stack([quat_wxyz_pose(pos, q) for pos, q in zip(p["base_positions"], p["base_orientation_wxyz"])])
        out.append(Problem(p["name"], n_arms, bases, scene_from_problem_dict(p),
                           np.asarray(p["start"], np.float32), np.asarray(p["goals"], np.float32).reshape(len(p["goals"]), -1),
                           int(p.get("source_index", 0))))
    return out


def problem_to_dict(p: Problem) -> dict:
    """Inverse of load_problem_file for generated problems (reference schema)."""
    def euler_from_axes(rec):
        r = np.stack([rec[3:6], rec[6:9], rec[9:12]], axis=1).astype(np.float64)
        theta = -math.asin(max(-1.0, min(1.0, r[2, 0])))
        rho = math.atan2(r[2, 1], r[2, 2])
        phi = math.atan2(r[1, 0], r[0, 0])
        return [rho, theta, phi]
    boxes = [{"name": f"box{i}", "position": c[0:3].tolist(), "orientation_euler_xyz": euler_from_axes(c),
              "half_extents": c[12:15].tolist()} for i, c in enumerate(p.scene.cuboids)]
    cyl = []
    for i, c in enumerate(p.scene.capsules):
        p1, v = c[0:3].astype(np.float64), c[3:6].astype(np.float64)
        length = float(np.linalg.norm(v))
        d = -v / length   # local +z points from p2 to p1
        theta = math.asin(max(-1.0, min(1.0, d[0])))
        rho = math.atan2(-d[1], d[2])
        cyl.append({"name": f"cylinder{i}", "position": (p1 + v / 2).tolist(),
                    "orientation_euler_xyz": [rho, theta, 0.0], "radius": float(c[6]), "length": length})
    sph = [{"name": f"sphere{i}", "position": s[0:3].tolist(), "radius": float(s[3])} for i, s in enumerate(p.scene.spheres)]
    pos = [b[[3, 7, 11]].tolist() for b in p.bases]
    return {"name": p.name, "valid": True, "source_index": p.seed, "base_positions": pos,
            "base_orientation_wxyz": [[1.0, 0.0, 0.0, 0.0]] * p.n_arms, "start": p.start.tolist(),
            "goals": p.goals.tolist(), "sphere": sph, "cylinder": cyl, "box": boxes}


# ------------------------------------------------------ scene generators ----
def default_bases(n_arms: int) -> np.ndarray:
    """Default base layout per arm count (4 and 5: the reference's tables)."""
    from . import geometry
    return geometry.default_bases(n_arms)


def _far_from_bases(xy, bases, clearance):
    return all(math.hypot(xy[0] - b[3], xy[1] - b[7]) > clearance for b in bases)


def scene_random_cuboids(rng: np.random.Generator, bases, n=16, half=(0.04, 0.12), lo=(-0.6, -0.6, 0.0),
                         hi=(0.6, 0.6, 1.2), base_clearance=0.3) -> Scene:
    """n axis-aligned cuboids, half-extents U[half]^3, centres U[lo, hi], kept
    `base_clearance` metres (in xy) away from every arm base."""
    cub = []
    while len(cub) < n:
        c = rng.uniform(lo, hi)
        if not _far_from_bases(c, bases, base_clearance):
            continue
        cub.append(cuboid_record(c, (0.0, 0.0, 0.0), rng.uniform(half[0], half[1], 3)))
    return Scene(cuboids=np.array(cub, np.float32))


def _random_obstacles(rng, n, bases, lo, hi, size, base_clearance, avoid=None):
    sph, cap, cub = [], [], []
    while len(sph) + len(cap) + len(cub) < n:
        c = rng.uniform(lo, hi)
        if not _far_from_bases(c, bases, base_clearance):
            continue
        if avoid is not None and avoid(c):
            continue
        kind = rng.integers(3)
        s = rng.uniform(size[0], size[1])
        if kind == 0:
            sph.append(sphere_record(c, s))
        elif kind == 1:
            cub.append(cuboid_record(c, (0.0, 0.0, rng.uniform(-math.pi, math.pi)), rng.uniform(size[0], size[1], 3)))
        else:
            length = rng.uniform(2 * size[0], 2 * size[1])
            cap.append(capsule_record(c, (rng.uniform(-math.pi, math.pi), rng.uniform(-math.pi / 2, math.pi / 2), 0.0),
                                      s, length))
    return sph, cap, cub


def bin_bases() -> np.ndarray:
    """4 arms at the corners (+-0.6, +-0.6, 0) of a 1.2 m table, facing the centre."""
    out = []
    for x, y in [(0.6, -0.6), (-0.6, -0.6), (-0.6, 0.6), (0.6, 0.6)]:
        out.append(yaw_pose(x, y, 0.0, math.atan2(-y, -x)))
    return np.stack(out)


def scene_bin(rng: np.random.Generator, bases, n_random=24) -> Scene:
    """Bin of 5 cuboids (1 cm walls, 40x40x20 cm) on a 1.2 m table, plus
    n_random spheres / cuboids / cylinders (capsules) of size 3-10 cm."""
    cub = [cuboid_record((0, 0, -0.01), (0, 0, 0), (0.6, 0.6, 0.01)),          # table top
           cuboid_record((0, 0, 0.005), (0, 0, 0), (0.2, 0.2, 0.005)),         # bin floor
           cuboid_record((0.195, 0, 0.1), (0, 0, 0), (0.005, 0.2, 0.1)),
           cuboid_record((-0.195, 0, 0.1), (0, 0, 0), (0.005, 0.2, 0.1)),
           cuboid_record((0, 0.195, 0.1), (0, 0, 0), (0.2, 0.005, 0.1)),
           cuboid_record((0, -0.195, 0.1), (0, 0, 0), (0.2, 0.005, 0.1))]
    def in_bin(c):
        return abs(c[0]) < 0.32 and abs(c[1]) < 0.32 and c[2] < 0.35
    sph, cap, cub2 = _random_obstacles(rng, n_random, bases, (-0.75, -0.75, 0.1), (0.75, 0.75, 1.0), (0.03, 0.10),
                                       0.35, in_bin)
    return Scene(np.array(sph, np.float32).reshape(-1, 4), np.array(cap, np.float32).reshape(-1, 8),
                 np.zeros((0, 8), np.float32), np.array(cub + cub2, np.float32).reshape(-1, 15))


def scene_shelf(rng: np.random.Generator, n_cylinders=15) -> Scene:
    """Shelf between the two rows of the 5-arm layout: 4 boards of 2 cm
    (1.0 m x 0.35 m) plus two side walls, with n_cylinders capsules
    (r 2-5 cm, h 8-20 cm) standing on the boards."""
    cx, cy = 0.3, 0.0
    levels = [0.25, 0.5, 0.75, 1.0]
    cub = [cuboid_record((cx, cy, z), (0, 0, 0), (0.5, 0.175, 0.01)) for z in levels]
    cub.append(cuboid_record((cx - 0.51, cy, 0.55), (0, 0, 0), (0.01, 0.175, 0.45)))
    cub.append(cuboid_record((cx + 0.51, cy, 0.55), (0, 0, 0), (0.01, 0.175, 0.45)))
    cap = []
    for _ in range(n_cylinders):
        z = levels[rng.integers(len(levels) - 1)] + 0.01
        r = rng.uniform(0.02, 0.05)
        h = rng.uniform(0.08, 0.20)
        x = rng.uniform(cx - 0.45, cx + 0.45)
        y = rng.uniform(cy - 0.13, cy + 0.13)
        cap.append(capsule_record((x, y, z + h / 2), (0.0, 0.0, 0.0), r, h))
    return Scene(capsules=np.array(cap, np.float32), cuboids=np.array(cub, np.float32))


SHELF_CENTRE = (0.3, 0.0)
SHELF_LEVELS = (0.25, 0.5, 0.75, 1.0)


def shelf_cells(margin: float = 0.03):
    """The 9 cells of scene_shelf (3 gaps between boards x 3 columns of the
    1 m width) as (low, high) corner pairs, shrunk by `margin` metres."""
    cx, cy = SHELF_CENTRE
    xs = np.linspace(cx - 0.49, cx + 0.49, 4)
    cells = []
    for z0, z1 in zip(SHELF_LEVELS[:-1], SHELF_LEVELS[1:]):
        for x0, x1 in zip(xs[:-1], xs[1:]):
            lo = np.array([x0, cy - 0.175, z0 + 0.01]) + margin
            hi = np.array([x1, cy + 0.175, z1 - 0.01]) - margin
            cells.append((lo, hi))
    return cells


def hand_positions(base: np.ndarray, q: np.ndarray) -> np.ndarray:
    """World position of the hand (mean centre of the 4 coarse spheres on
    joint 7) of one arm with base pose `base` (12 floats, row-major 3x4) for
    configurations q [n, 7], by the kinematic chain T_j = T_(j-1) FT_j Rz(q_j)
    of the sphere model (float64)."""
    from . import geometry
    ft = geometry.table("kJointTf").reshape(8, 3, 4).astype(np.float64)
    sph = geometry.table("kCoarseSphere").reshape(11, 4).astype(np.float64)
    q = np.asarray(q, np.float64).reshape(-1, 7)
    n = len(q)
    B = np.asarray(base, np.float64).reshape(3, 4)
    col = [np.broadcast_to(B[:, k], (n, 3)) for k in range(4)]  # columns of the 3x4 frame
    for j in range(1, 8):
        F = ft[j]
        c, s = np.cos(q[:, j - 1])[:, None], np.sin(q[:, j - 1])[:, None]
        # M = F Rz(q): columns F0 c + F1 s, F1 c - F0 s, F2, F3; new frame = T M
        def mul(v):  # T applied to a direction v (3 rows of F)
            return col[0] * v[0] + col[1] * v[1] + col[2] * v[2]
        t0, t1, t2 = mul(F[:, 0]), mul(F[:, 1]), mul(F[:, 2])
        col = [t0 * c + t1 * s, t1 * c - t0 * s, t2, mul(F[:, 3]) + col[3]]
    off = sph[7:11, :3]
    return sum(col[0] * o[0] + col[1] * o[1] + col[2] * o[2] + col[3] for o in off) / 4.0


_CELL_POOLS = {}


def _cell_pools(base, cells, count=512, batch=65536, rounds=8):
    """Per cell, up to `count` arm configurations (uniform in the joint
    limits) whose hand lies inside the cell.  They depend on the base pose
    only, so they are drawn once per pose from a fixed stream and cached."""
    key = (np.asarray(base, np.float32).tobytes(), count)
    if key in _CELL_POOLS:
        return _CELL_POOLS[key]
    rng = np.random.default_rng(0x5E1F)
    found = [[] for _ in cells]
    for _ in range(rounds):
        u = np.maximum(rng.uniform(0.0, 1.0, (batch, DOF_PER_ARM)), 1e-7)
        q = (u * SCALE + LOW).astype(np.float32)
        p = hand_positions(base, q)
        for k, (lo, hi) in enumerate(cells):
            if len(found[k]) < count:
                found[k].extend(q[((p > lo) & (p < hi)).all(axis=1)][:count - len(found[k])])
        if all(len(f) >= count for f in found):
            break
    _CELL_POOLS[key] = [np.array(f, np.float32).reshape(-1, DOF_PER_ARM) for f in found]
    return _CELL_POOLS[key]


def reachable(rng: np.random.Generator, scene: Scene, base: np.ndarray, q: np.ndarray, checker: "Checker",
              reach: float = 2.0, probes: int = 32, points: int = 16, min_free: int = 3) -> np.ndarray:
    """Which one-arm configurations q [n, 7] can be left in a straight line:
    of `probes` segments from q to points uniform in the joint-space ball of
    radius `reach` around it (clipped to the joint limits), at least
    `min_free` are collision-free for the arm alone (environment + self,
    checked at `points` evenly spaced points).  A roadmap links a node to
    the nodes within its range by such segments, so a configuration deep in
    a shelf cell that sees almost none of that ball (the hand walled in by
    boards and cylinders) is a start or goal no roadmap connects."""
    q = np.asarray(q, np.float32).reshape(-1, DOF_PER_ARM)
    n = len(q)
    if n == 0:
        return np.zeros(0, bool)
    d = rng.normal(size=(n, probes, DOF_PER_ARM))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    r = reach * rng.uniform(0.0, 1.0, (n, probes, 1)) ** (1.0 / DOF_PER_ARM)
    x = np.clip(q[:, None, :] + d * r, LOW, LOW + SCALE)
    t = (np.arange(1, points + 1) / points)[None, None, :, None]
    pts = (q[:, None, None, :] + (x - q[:, None, :])[:, :, None, :] * t).astype(np.float32)
    coll = checker(scene, 1, np.asarray(base)[None], pts.reshape(-1, DOF_PER_ARM)).reshape(n, probes, points)
    return (~coll.any(axis=2)).sum(axis=1) >= min_free


def shelf_cell_configs(rng: np.random.Generator, scene: Scene, bases: np.ndarray, count: int, checker: "Checker",
                       tries: int = 64, batch: int = 2048) -> np.ndarray:
    """`count` collision-free multi-arm configurations with every arm's hand
    inside a shelf cell, the arms in distinct cells (configs[2] of the
    benchmark: end-effectors reaching into random shelf cells).  Per arm and
    cell a pool of in-cell arm configurations is drawn by rejection on the
    hand position; those that are free for the arm alone and reachable (see
    `reachable`) form the cell's pool;
    composites of pool members over random cell assignments are then checked
    whole."""
    n_arms = len(bases)
    cells = shelf_cells()
    pools = []
    for a in range(n_arms):
        row = []
        for pl in _cell_pools(bases[a], cells):
            if len(pl):
                pl = pl[~checker(scene, 1, bases[a:a + 1], pl)]  # free on its own (environment + self)
                pl = pl[reachable(rng, scene, bases[a], pl, checker)]
            row.append(pl)
        pools.append(row)
    reach = [[k for k, pl in enumerate(pools[a]) if len(pl) >= 2] for a in range(n_arms)]
    if any(not r for r in reach):
        raise RuntimeError("an arm reaches no shelf cell")
    out = []
    for _ in range(tries):
        cand = np.zeros((batch, DOF_PER_ARM * n_arms), np.float32)
        placed = np.ones(batch, bool)
        for i in range(batch):
            used = set()
            for a in rng.permutation(n_arms):
                free = [k for k in reach[a] if k not in used]
                if not free:  # every cell this arm reaches is taken: no composite
                    placed[i] = False
                    break
                k = free[rng.integers(len(free))]
                used.add(k)
                pl = pools[a][k]
                cand[i, DOF_PER_ARM * a:DOF_PER_ARM * (a + 1)] = pl[rng.integers(len(pl))]
        cand = cand[placed]
        ok = ~checker(scene, n_arms, bases, cand) if len(cand) else np.zeros(0, bool)
        out.extend(cand[ok])
        if len(out) >= count:
            return np.stack(out[:count]).astype(np.float32)
    raise RuntimeError(f"could not sample {count} in-cell configurations ({len(out)} found)")


def scene_posts(rng: np.random.Generator, bases, n_posts=10, n_boxes=10) -> Scene:
    """Work cell with the reference's z-aligned obstacle types
    (utils.cuh:165-181, 208-228): vertical posts as z-capsules (r 2-6 cm,
    h 10-60 cm) and yaw-rotated boxes as z-cuboids (half-extents 3-12 cm),
    plus a table top and a few spheres."""
    zcap, zcub = [], []
    while len(zcap) < n_posts:
        x, y = rng.uniform(-0.7, 0.7, 2)
        if not _far_from_bases((x, y), bases, 0.3):
            continue
        z0, h, r = rng.uniform(0.0, 0.5), rng.uniform(0.1, 0.6), rng.uniform(0.02, 0.06)
        zcap.append([x, y, z0, 0.0, 0.0, h, r, 1.0 / (h * h)])
    while len(zcub) < n_boxes:
        c = rng.uniform((-0.7, -0.7, 0.1), (0.7, 0.7, 1.0))
        if not _far_from_bases(c, bases, 0.3):
            continue
        yaw = rng.uniform(-math.pi, math.pi)
        cz, sz = math.cos(yaw), math.sin(yaw)
        hx, hy, hz = rng.uniform(0.03, 0.12, 3)
        zcub.append([c[0], c[1], c[2], cz, sz, 0.0, -sz, cz, 0.0, 0.0, 0.0, 1.0, hx, hy, hz])
    sph = [sphere_record(rng.uniform((-0.6, -0.6, 0.3), (0.6, 0.6, 1.0)), rng.uniform(0.03, 0.08)) for _ in range(3)]
    table = [cuboid_record((0, 0, -0.01), (0, 0, 0), (0.8, 0.8, 0.01))]
    return Scene(spheres=np.array(sph, np.float32), z_capsules=np.array(zcap, np.float32),
                 cuboids=np.array(table, np.float32), z_cuboids=np.array(zcub, np.float32))


def scene_table3(rng: np.random.Generator, bases, n_random=12) -> Scene:
    """Round 3-arm work cell: a table top plus n_random mixed obstacles."""
    cub = [cuboid_record((0, 0, -0.01), (0, 0, 0), (0.8, 0.8, 0.01))]
    sph, cap, cub2 = _random_obstacles(rng, n_random, bases, (-0.6, -0.6, 0.1), (0.6, 0.6, 1.0), (0.03, 0.10), 0.35)
    return Scene(np.array(sph, np.float32).reshape(-1, 4), np.array(cap, np.float32).reshape(-1, 8),
                 np.zeros((0, 8), np.float32), np.array(cub + cub2, np.float32).reshape(-1, 15))


Checker = Callable[[Scene, int, np.ndarray, np.ndarray], np.ndarray]
"""checker(scene, n_arms, bases, cfgs[n, 7*n_arms]) -> bool[n] (True = colliding),
for single configurations under the full multi-arm composite check."""


def sample_free_configs(rng: np.random.Generator, scene: Scene, n_arms: int, bases: np.ndarray, count: int,
                        checker: Checker, batch: int = 4096, max_rounds: int = 400) -> np.ndarray:
    """`count` collision-free configurations by rejection sampling, uniform in
    the joint limits."""
    dim = DOF_PER_ARM * n_arms
    found = []
    for _ in range(max_rounds):
        u = rng.uniform(0.0, 1.0, (batch, dim)).astype(np.float32)
        u = np.maximum(u, np.float32(1e-7))
        q = u * np.tile(SCALE, n_arms) + np.tile(LOW, n_arms)
        ok = ~checker(scene, n_arms, bases, q)
        found.extend(q[ok])
        if len(found) >= count:
            return np.stack(found[:count]).astype(np.float32)
    raise RuntimeError(f"could not sample {count} free configurations ({len(found)} found)")


def make_problem(kind: str, seed: int, checker: Checker, n_goals: int = 1) -> Problem:
    """Seeded synthetic problem.  kind: 'plain2' (2 arms, 16 random cuboids),
    'table3' (3 arms), 'bin4' (4 arms around a bin), 'shelf5' (5 arms, shelf),
    'shelf5cell' (the shelf with every hand inside a shelf cell at the start
    and the goals), 'posts3' (3 arms among z-aligned posts and boxes)."""
    rng = np.random.default_rng([seed, {"plain2": 2, "table3": 3, "bin4": 4, "shelf5": 5, "shelf5cell": 6,
                                        "posts3": 7}[kind]])
    if kind == "shelf5cell":
        # a draw whose cylinders leave an arm no free cell (or no free
        # composite) is replaced by the next scene of the same stream
        n_arms, bases = 5, default_bases(5)
        for _ in range(16):
            scene = scene_shelf(rng)
            try:
                cfgs = shelf_cell_configs(rng, scene, bases, 1 + n_goals, checker)
            except RuntimeError:
                continue
            return Problem(f"{kind}_{seed}", n_arms, bases, scene, cfgs[0], cfgs[1:], seed)
        raise RuntimeError(f"{kind} seed {seed}: no scene with free in-cell configurations in 16 draws")
    if kind == "plain2":
        n_arms, bases = 2, default_bases(2)
        scene = scene_random_cuboids(rng, bases)
    elif kind == "table3":
        n_arms, bases = 3, default_bases(3)
        scene = scene_table3(rng, bases)
    elif kind == "bin4":
        n_arms, bases = 4, bin_bases()
        scene = scene_bin(rng, bases)
    elif kind == "shelf5":
        n_arms, bases = 5, default_bases(5)
        scene = scene_shelf(rng)
    elif kind == "posts3":
        n_arms, bases = 3, default_bases(3)
        scene = scene_posts(rng, bases)
    else:
        raise ValueError(kind)
    cfgs = sample_free_configs(rng, scene, n_arms, bases, 1 + n_goals, checker)
    return Problem(f"{kind}_{seed}", n_arms, bases, scene, cfgs[0], cfgs[1:], seed)


def scene_records_equal(a: Scene, b: Scene) -> bool:
    return all(np.array_equal(getattr(a, k), getattr(b, k))
               for k in ("spheres", "capsules", "z_capsules", "cuboids", "z_cuboids"))

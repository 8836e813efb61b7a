"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
EPS_CONTACT = 1e-5  # metres: samples whose verdict a clearance this small could flip are exempt

MODE_FULL, MODE_DRRT, MODE_SINGLE = 0, 1, 2

_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        L.orc_sphere_centres.argtypes = [ctypes.c_int, P, P, P, P]
        L.orc_arm_frames.argtypes = [P, P, P]
        L.orc_env_clearance.argtypes = [P, P, ctypes.c_double]
        L.orc_env_clearance.restype = ctypes.c_double
        L.orc_check_config.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_double,
                                       P, P, P, P]
        L.orc_check_edges.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int,
                                      P, P, ctypes.c_double, P, P, P, P, P, P]
        L.orc_check_edges_f32.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                          ctypes.c_int, P, P, P, ctypes.c_int]
        L.orc_check_edges_ref_schedule.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                                   ctypes.c_int, P, P, P, P, P]
        L.orc_ref_thread_shares.argtypes = [P, ctypes.c_int, P]
        L.orc_tables.argtypes = [P] * 11
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def sphere_centres(bases, q):
    """FP64 centres of one configuration: (coarse[A,11,3], fine[A,59,3])."""
    bases = _f32(bases).reshape(-1, 12)
    A = len(bases)
    q = _f64(q).reshape(-1)
    assert q.size == 7 * A
    co = np.zeros((A, 11, 3))
    fi = np.zeros((A, 59, 3))
    lib().orc_sphere_centres(A, bases.ctypes.data, q.ctypes.data, co.ctypes.data, fi.ctypes.data)
    return co, fi


def arm_frames(base, q):
    base = _f32(base).reshape(12)
    q = _f64(q).reshape(7)
    out = np.zeros((8, 3, 4))
    lib().orc_arm_frames(base.ctypes.data, q.ctypes.data, out.ctypes.data)
    return out


def env_clearance(scene, p, r):
    cs = scene.as_c()
    p = _f64(p).reshape(3)
    return lib().orc_env_clearance(ctypes.byref(cs), p.ctypes.data, float(r))


def check_config(scene, mode, bases, q, arm=0, eps=EPS_CONTACT):
    """(nominal, lo, hi, min_abs_clearance) for one configuration (FP64)."""
    bases = _f32(bases).reshape(-1, 12)
    q = _f64(q).reshape(-1)
    cs = scene.as_c()
    nom, lo, hi = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    m = ctypes.c_double()
    lib().orc_check_config(ctypes.byref(cs), mode, len(bases), arm, bases.ctypes.data, q.ctypes.data, eps,
                           ctypes.byref(nom), ctypes.byref(lo), ctypes.byref(hi), ctypes.byref(m))
    return bool(nom.value), bool(lo.value), bool(hi.value), m.value


def check_edges(scene, mode, bases, G, frm, to, arm=0, eps=EPS_CONTACT):
    """FP64 edge verdicts: dict with nom/lo/hi (bool[n]) and mask_nom/lo/hi (uint64[n])."""
    bases = _f32(bases).reshape(-1, 12)
    frm, to = _f32(frm), _f32(to)
    n = frm.shape[0]
    cs = scene.as_c()
    out = {k: np.zeros(n, np.uint8) for k in ("nom", "lo", "hi")}
    out.update({k: np.zeros(n, np.uint64) for k in ("mask_nom", "mask_lo", "mask_hi")})
    lib().orc_check_edges(ctypes.byref(cs), mode, len(bases), arm, bases.ctypes.data, G, n, frm.ctypes.data,
                          to.ctypes.data, eps, out["nom"].ctypes.data, out["lo"].ctypes.data, out["hi"].ctypes.data,
                          out["mask_nom"].ctypes.data, out["mask_lo"].ctypes.data, out["mask_hi"].ctypes.data)
    for k in ("nom", "lo", "hi"):
        out[k] = out[k].astype(bool)
    return out


def check_edges_plain_ranges(scene, mode, bases, G, frm, to, arm=0, eps=EPS_CONTACT):
    """check_edges with the multi-arm fine self ranges read the single-arm
    way (not the reference's reading, panda_four.cuh:390-401): only for
    finding samples where that reading changes the verdict."""
    lib().orc_set_plain_multi_arm_ranges(1)
    try:
        return check_edges(scene, mode, bases, G, frm, to, arm=arm, eps=eps)
    finally:
        lib().orc_set_plain_multi_arm_ranges(0)


def check_edges_f32(scene, mode, bases, G, frm, to, arm=0, threads=1):
    """FP32 mirror of the reference checker with early exits (CPU baseline)."""
    bases = _f32(bases).reshape(-1, 12)
    frm, to = _f32(frm), _f32(to)
    n = frm.shape[0]
    cs = scene.as_c()
    out = np.zeros(n, np.uint8)
    lib().orc_check_edges_f32(ctypes.byref(cs), mode, len(bases), arm, bases.ctypes.data, G, n, frm.ctypes.data,
                              to.ctypes.data, out.ctypes.data, int(threads))
    return out.astype(bool)


def ref_thread_shares(counts):
    """Threads of the reference's fine stages per sample for per-sample sphere
    counts (utils.cuh:477-498): int32[G]."""
    c = np.ascontiguousarray(counts, np.int32)
    out = np.zeros(len(c), np.int32)
    lib().orc_ref_thread_shares(c.ctypes.data, len(c), out.ctypes.data)
    return out


def check_edges_ref_schedule(scene, mode, bases, G, frm, to, arm=0):
    """FP64 nominal edge verdicts with and without the reference's fine-stage
    thread schedule: dict sched (bool[n]), all (bool[n]), dropped (uint64[n]:
    colliding samples whose fine stages got no thread)."""
    bases = _f32(bases).reshape(-1, 12)
    frm, to = _f32(frm), _f32(to)
    n = frm.shape[0]
    cs = scene.as_c()
    sched, full = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    dropped = np.zeros(n, np.uint64)
    lib().orc_check_edges_ref_schedule(ctypes.byref(cs), mode, len(bases), arm, bases.ctypes.data, G, n,
                                       frm.ctypes.data, to.ctypes.data, sched.ctypes.data, full.ctypes.data,
                                       dropped.ctypes.data)
    return {"sched": sched.astype(bool), "all": full.astype(bool), "dropped": dropped}


def config_checker(scene, n_arms, bases, cfgs):
    """scenes.Checker backed by the FP32 mirror: True = colliding (G = 1 edges from a config to itself)."""
    cfgs = _f32(cfgs).reshape(len(cfgs), -1)
    mode = MODE_SINGLE if n_arms == 1 else MODE_FULL  # one arm: environment + self
    return check_edges_f32(scene, mode, bases, 1, cfgs, cfgs, threads=os.cpu_count() or 1)


def tables():
    out = dict(joint_tf=np.zeros((8, 12), np.float32), coarse=np.zeros((11, 4), np.float32),
               coarse_joint=np.zeros(11, np.int32), fine=np.zeros((59, 4), np.float32),
               fine_joint=np.zeros(59, np.int32), coarse_ranges=np.zeros((5, 3), np.int32),
               fine_ranges=np.zeros((25, 3), np.int32), scale=np.zeros(7, np.float32), low=np.zeros(7, np.float32),
               base4=np.zeros((4, 12), np.float32), base5=np.zeros((5, 12), np.float32))
    lib().orc_tables(*[v.ctypes.data for v in out.values()])
    return out

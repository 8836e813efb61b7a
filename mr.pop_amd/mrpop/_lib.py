"""ctypes binding of libmrpop.so (include/mrpop_c.h).

The library is the product: it must be built (mr.pop_amd/csrc/Makefile) and a
GPU must be present for any compute call; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import weakref
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent          # mr.pop_amd/
LIB_PATH = PKG / "libmrpop.so"
HEADER = PKG.parent / "include" / "mrpop_c.h"

OK, ERR_HIP, ERR_INVALID, ERR_BUDGET, ERR_TIMEOUT, ERR_NO_DEVICE, ERR_STATE, ERR_INTERNAL = range(8)
MODE_FULL, MODE_DRRT, MODE_SINGLE = 0, 1, 2

P = ctypes.c_void_p
I = ctypes.c_int


class Settings(ctypes.Structure):
    """mrpop_settings == pop_settings (src/planning/pop_settings.hh:3-32)."""
    _fields_ = [("max_samples", I), ("max_iters", I), ("optimize_iters", I), ("num_new_configs", I),
                ("granularity", I), ("range", ctypes.c_float), ("balance", I), ("tree_ratio", ctypes.c_float),
                ("dynamic_domain", I), ("dd_alpha", ctypes.c_float), ("dd_radius", ctypes.c_float),
                ("dd_min_radius", ctypes.c_float), ("roadmap_reassign", I), ("rrtc_iter", I), ("resample_iter", I),
                ("cost_bound_resample", I), ("path_simplify", I), ("phs", I), ("simplify_round", I),
                ("bspline_step", I), ("bspline_min_change", ctypes.c_float), ("perturb_step", I),
                ("perturb_empty_step", I)]


class SolveOptions(ctypes.Structure):
    """mrpop_solve_options (include/mrpop_c.h)."""
    _fields_ = [("block_per_robot", I), ("time_budget_s", ctypes.c_double), ("roadmap_capacity", I),
                ("tree_capacity", I), ("neighbour_capacity", I), ("neighbour_candidates", I), ("rounds_per_step", I),
                ("probe_samples", I), ("graph_rounds", I), ("arm_split", I), ("biased_iter", I),
                ("roadmap_cache", I)]


class CProblem(ctypes.Structure):
    _fields_ = [("n_arms", I), ("bases", P), ("scene", P), ("start", P), ("n_goals", I), ("goals", P),
                ("seed", ctypes.c_uint64)]


U64 = ctypes.c_uint64


class ResultInfo(ctypes.Structure):
    _fields_ = [("solved", I), ("cost", ctypes.c_float), ("path_length", I), ("start_tree_size", I),
                ("goal_tree_size", I), ("iters", I), ("wall_ns", U64), ("kernel_ns", U64), ("copy_ns", U64),
                ("ttfs_ns", U64), ("n_seed_paths", I), ("passes", I), ("rounds", I), ("roadmap_rounds", I),
                ("roadmap_size", I * 5), ("status", I), ("union_find_bound", I), ("edge_table_full", I),
                ("connect_skipped", I), ("simplify_dropped", I), ("roadmap_cached", I * 5)]


class MrpopError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mrpop error {code}: {msg}")
        self.code = code


# name -> (restype, argtypes)
SIGNATURES = {
    "mrpop_version": (ctypes.c_char_p, []),
    "mrpop_settings_default": (I, [P]),
    "mrpop_default_bases": (I, [I, P]),
    "mrpop_create": (I, [I, P]),
    "mrpop_destroy": (I, [P]),
    "mrpop_last_error": (ctypes.c_char_p, [P]),
    "mrpop_stream": (P, [P]),
    "mrpop_world_create": (I, [P, P, I, P, P]),
    "mrpop_world_destroy": (I, [P]),
    "mrpop_check_edges": (I, [P, P, I, I, I, I, P, P, P, P]),
    "mrpop_check_edges_device": (I, [P, P, I, I, I, I, P, P, P, P, P]),
    "mrpop_check_edges_noalloc": (I, [P, P, I, I, I, I, P, P, P, P]),
    "mrpop_sphere_centres": (I, [P, P, I, P, P, P]),
    "mrpop_solve_options_default": (I, [P]),
    "mrpop_solve": (I, [P, P, P, P, P]),
    "mrpop_solve_batch": (I, [P, I, P, P, P, P]),
    "mrpop_solution_info": (I, [P, P]),
    "mrpop_solution_path": (I, [P, I, P, I, P, P]),
    "mrpop_solution_improvements": (I, [P, P, I, P]),
    "mrpop_solution_free": (I, [P]),
    "mrpop_server_create": (I, [P, ctypes.c_size_t, I, P]),
    "mrpop_server_destroy": (I, [P]),
    "mrpop_server_submit": (I, [P, P, P, P, P]),
    "mrpop_server_step": (I, [P, P, P]),
    "mrpop_server_poll": (I, [P, U64, P, P]),
    "mrpop_server_cancel": (I, [P, U64]),
    "mrpop_server_bytes": (ctypes.c_size_t, [P]),
}

_lib = None


def lib():
    """Load libmrpop.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C mr.pop_amd/csrc`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def settings_default() -> Settings:
    s = Settings()
    lib().mrpop_settings_default(ctypes.byref(s))
    return s


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Context:
    """A device context (owns a HIP stream and workspace)."""

    def __init__(self, device: int = 0):
        self._h = P()
        self._children = weakref.WeakSet()   # worlds etc. are released before the context
        rc = lib().mrpop_create(device, ctypes.byref(self._h))
        if rc != OK:
            raise MrpopError(rc, "mrpop_create failed (no GPU?)")

    def check(self, rc):
        if rc != OK:
            raise MrpopError(rc, lib().mrpop_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return lib().mrpop_stream(self._h) or 0

    def close(self):
        if self._h:
            for child in list(self._children):
                child.close()
            lib().mrpop_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def world(self, scene, n_arms, bases=None) -> "World":
        return World(self, scene, n_arms, bases)


class World:
    """A scene plus arm layout uploaded to the device."""

    def __init__(self, ctx: Context, scene, n_arms: int, bases=None):
        self.ctx = ctx
        self.n_arms = n_arms
        self._scene = scene
        self._cs = scene.as_c()
        self._bases = None if bases is None else _f32(bases).reshape(n_arms, 12)
        self._h = P()
        ctx.check(lib().mrpop_world_create(ctx.handle, ctypes.byref(self._cs), n_arms,
                                           None if self._bases is None else self._bases.ctypes.data,
                                           ctypes.byref(self._h)))
        ctx._children.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().mrpop_world_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check_edges(self, mode, granularity, frm, to, arm=0, with_mask=False, noalloc=False):
        """Host-buffer edge check: bool[n] (and uint64[n] sample masks).
        noalloc: stage through the library's fixed per-device array
        (mrpop_check_edges_noalloc) instead of context buffers."""
        frm, to = _f32(frm), _f32(to)
        n = frm.shape[0]
        out = np.zeros(n, np.uint8)
        mask = np.zeros(n, np.uint64) if with_mask else None
        fn = lib().mrpop_check_edges_noalloc if noalloc else lib().mrpop_check_edges
        self.ctx.check(fn(self.ctx.handle, self._h, mode, arm, granularity, n, frm.ctypes.data, to.ctypes.data,
                          out.ctypes.data, None if mask is None else mask.ctypes.data))
        return (out.astype(bool), mask) if with_mask else out.astype(bool)

    def check_edges_device(self, mode, granularity, n, d_from, d_to, d_out, d_mask=None, stream=None, arm=0):
        """Device-pointer edge check (ints are raw device addresses); enqueued, not synchronised."""
        self.ctx.check(lib().mrpop_check_edges_device(self.ctx.handle, self._h, mode, arm, granularity, n,
                                                      d_from, d_to, d_out, d_mask, stream))

    def sphere_centres(self, cfgs):
        cfgs = _f32(cfgs).reshape(-1, 7 * self.n_arms)
        n = len(cfgs)
        co = np.zeros((n, self.n_arms, 11, 3), np.float32)
        fi = np.zeros((n, self.n_arms, 59, 3), np.float32)
        self.ctx.check(lib().mrpop_sphere_centres(self.ctx.handle, self._h, n, cfgs.ctypes.data, co.ctypes.data,
                                                  fi.ctypes.data))
        return co, fi

    def config_collides(self, cfgs):
        """Full multi-arm verdict of single configurations (G = 1 edges cfg -> cfg);
        a one-arm world uses the single-arm check (environment + self)."""
        cfgs = _f32(cfgs).reshape(-1, 7 * self.n_arms)
        return self.check_edges(MODE_SINGLE if self.n_arms == 1 else MODE_FULL, 1, cfgs, cfgs)


def gpu_checker(ctx: Context):
    """scenes.Checker backed by the GPU edge checker."""
    cache = {}

    def checker(scene, n_arms, bases, cfgs):
        key = (id(scene), n_arms, None if bases is None else _f32(bases).tobytes())
        if key not in cache:
            cache.clear()
            cache[key] = World(ctx, scene, n_arms, bases)
        return cache[key].config_collides(cfgs)
    return checker


# --------------------------------------------------------------- planning --
def solve_options_default(**kw) -> SolveOptions:
    o = SolveOptions()
    lib().mrpop_solve_options_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def settings(**kw) -> Settings:
    """pop_settings defaults with overrides."""
    s = settings_default()
    for k, v in kw.items():
        if not hasattr(s, k):
            raise AttributeError(k)
        setattr(s, k, v)
    return s


class _ProblemRef:
    """Keeps the numpy arrays behind a CProblem alive."""

    def __init__(self, problem):
        self.p = problem
        self.scene_c = problem.scene.as_c()
        self.bases = _f32(problem.bases).reshape(-1)
        self.start = _f32(problem.start).reshape(-1)
        self.goals = _f32(problem.goals).reshape(-1)
        self.c = CProblem(problem.n_arms, self.bases.ctypes.data, ctypes.addressof(self.scene_c),
                          self.start.ctypes.data, len(problem.goals), self.goals.ctypes.data, int(problem.seed))


class Solution:
    """Result of a solve: PlannerResult fields plus timing."""

    def __init__(self, handle, dim):
        self.dim = dim
        info = ResultInfo()
        lib().mrpop_solution_info(handle, ctypes.byref(info))
        self.info = {f: getattr(info, f) for f, _ in ResultInfo._fields_ if f not in ("roadmap_size", "roadmap_cached")}
        self.info["roadmap_size"] = list(info.roadmap_size)
        self.info["roadmap_cached"] = list(info.roadmap_cached)
        self.solved = bool(info.solved)
        self.cost = float(info.cost)
        self.path = self._path(handle, -1)
        self.seed_paths = [self._path(handle, k) for k in range(info.n_seed_paths)]
        n = ctypes.c_int()
        lib().mrpop_solution_improvements(handle, None, 0, ctypes.byref(n))
        arr = (U64 * max(1, n.value))()
        lib().mrpop_solution_improvements(handle, arr, n.value, ctypes.byref(n))
        self.improve_s = [arr[i] * 1e-9 for i in range(n.value)]
        lib().mrpop_solution_free(handle)

    def _path(self, handle, which):
        n = ctypes.c_int()
        c = ctypes.c_float()
        lib().mrpop_solution_path(handle, which, None, 0, ctypes.byref(n), ctypes.byref(c))
        out = np.zeros((n.value, self.dim), np.float32)
        lib().mrpop_solution_path(handle, which, out.ctypes.data, n.value, ctypes.byref(n), ctypes.byref(c))
        return out if which < 0 else (float(c.value), out)

    @property
    def ttfs_s(self):
        return self.info["ttfs_ns"] * 1e-9

    @property
    def wall_s(self):
        return self.info["wall_ns"] * 1e-9


def slot_bytes(ctx: Context, n_arms: int, settings=None, options=None) -> int:
    """Device bytes of one problem (what a server's HBM budget admits against); 0 if rejected."""
    fn = lib().mrpop_diag_slot_bytes
    fn.restype = ctypes.c_size_t
    fn.argtypes = [P, I, P, P]
    return int(fn(ctx.handle, n_arms, None if settings is None else ctypes.byref(settings),
                  None if options is None else ctypes.byref(options)))


def solve(ctx: Context, problem, settings=None, options=None) -> Solution:
    ref = _ProblemRef(problem)
    out = P()
    ctx.check(lib().mrpop_solve(ctx.handle, ctypes.byref(ref.c), None if settings is None else ctypes.byref(settings),
                                None if options is None else ctypes.byref(options), ctypes.byref(out)))
    return Solution(out, problem.dim)


def solve_batch(ctx: Context, problems, settings=None, options=None):
    refs = [_ProblemRef(p) for p in problems]
    n = len(refs)
    arr = (CProblem * n)(*[r.c for r in refs])
    st = None if settings is None else (Settings * n)(*settings)
    op = None if options is None else (SolveOptions * n)(*options)
    outs = (P * n)()
    ctx.check(lib().mrpop_solve_batch(ctx.handle, n, arr, st, op, outs))
    return [Solution(outs[i], problems[i].dim) for i in range(n)]


class Server:
    """Continuous batching: submit problems at any time, step, poll."""

    def __init__(self, ctx: Context, hbm_budget_bytes=0, max_active=0):
        self.ctx = ctx
        self._h = P()
        self._refs = {}
        ctx.check(lib().mrpop_server_create(ctx.handle, int(hbm_budget_bytes), int(max_active), ctypes.byref(self._h)))
        ctx._children.add(self)

    def submit(self, problem, settings=None, options=None) -> int:
        ref = _ProblemRef(problem)
        t = U64()
        self.ctx.check(lib().mrpop_server_submit(self._h, ctypes.byref(ref.c),
                                                 None if settings is None else ctypes.byref(settings),
                                                 None if options is None else ctypes.byref(options), ctypes.byref(t)))
        self._refs[t.value] = (ref, problem.dim)
        return t.value

    def step(self):
        a, w = I(), I()
        self.ctx.check(lib().mrpop_server_step(self._h, ctypes.byref(a), ctypes.byref(w)))
        return a.value, w.value

    def poll(self, ticket, snapshot=False):
        st = I()
        out = P()
        self.ctx.check(lib().mrpop_server_poll(self._h, ticket, ctypes.byref(st), ctypes.byref(out) if snapshot else None))
        sol = Solution(out, self._refs[ticket][1]) if snapshot and out else None
        return st.value, sol

    def cancel(self, ticket):
        self.ctx.check(lib().mrpop_server_cancel(self._h, ticket))

    def bytes(self):
        return lib().mrpop_server_bytes(self._h)

    def close(self):
        if self._h:
            lib().mrpop_server_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

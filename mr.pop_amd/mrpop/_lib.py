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

OK, ERR_HIP, ERR_INVALID, ERR_BUDGET, ERR_TIMEOUT, ERR_NO_DEVICE, ERR_STATE = range(7)
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
    "mrpop_sphere_centres": (I, [P, P, I, P, P, P]),
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

    def check_edges(self, mode, granularity, frm, to, arm=0, with_mask=False):
        """Host-buffer edge check: bool[n] (and uint64[n] sample masks)."""
        frm, to = _f32(frm), _f32(to)
        n = frm.shape[0]
        out = np.zeros(n, np.uint8)
        mask = np.zeros(n, np.uint64) if with_mask else None
        self.ctx.check(lib().mrpop_check_edges(self.ctx.handle, self._h, mode, arm, granularity, n,
                                               frm.ctypes.data, to.ctypes.data, out.ctypes.data,
                                               None if mask is None else mask.ctypes.data))
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
        """Full multi-arm verdict of single configurations (G = 1 edges cfg -> cfg)."""
        cfgs = _f32(cfgs).reshape(-1, 7 * self.n_arms)
        return self.check_edges(MODE_FULL, 1, cfgs, cfgs)


def gpu_checker(ctx: Context):
    """scenes.Checker backed by the GPU edge checker."""
    cache = {}

    def checker(scene, n_arms, bases, cfgs):
        key = (id(scene), n_arms)
        if key not in cache:
            cache.clear()
            cache[key] = World(ctx, scene, n_arms, bases)
        return cache[key].config_collides(cfgs)
    return checker

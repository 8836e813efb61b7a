"""CPU restatement of the planner's path simplification stages.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

Reference: path_simplification.cu:41-238, as the build runs it.
  * Smoothing stages (bspline_step of them): every odd waypoint i < N-1 gets
    the candidate ((p[i-1]+p[i])/2 + (p[i]+p[i+1])/2)/2; a candidate that
    moves the waypoint by at least min_change is kept if both half-edges
    p[i-1] -> mid and p[i+1] -> mid, each split into ceil(d / range) chunks,
    are free.  All candidates of a stage are computed from the path as it
    stood before the stage.
  * Shortcut stage: every (i, j = i+b+1), b = 1..15, is checked in
    ceil((d - 0.001) / range) chunks; a DP over the free shortcuts (and
    every i -> i+1) gives the cheapest chain; if it beats the best cost the
    skipped waypoints are overwritten with copies of the waypoint before
    them, so the path keeps its length.
Chunk edges of a stage are numbered in candidate order; a candidate whose
chunks would pass the edge capacity is treated as colliding.  Chunk
configurations advance by delta = (to - from) / n per chunk, as in the
reference.  Float32 arithmetic as on the device (search.fma32 sums).
"""
from __future__ import annotations

import math

import numpy as np

from .search import sqdist, sqrt32

FLT_MAX = np.float32(3.402823466e38)


def l2(a, b) -> np.float32:
    return sqrt32(sqdist(np.asarray(a, np.float32)[None], np.asarray(b, np.float32))[0])


def chunks(frm, to, n: int):
    """The n chunk edges frm -> to: (from, to) arrays [n][D]."""
    frm = np.asarray(frm, np.float32)
    delta = ((np.asarray(to, np.float32) - frm) / np.float32(n)).astype(np.float32)
    f, t = [], []
    cur = frm.copy()
    for _ in range(n):
        f.append(cur.copy())
        cur = (cur + delta).astype(np.float32)
        t.append(cur.copy())
    return np.array(f), np.array(t)


def stage_edges(path: np.ndarray, smooth: bool, rng: float, min_change: float, cap: int):
    """Candidates and chunk edges of one stage.  Returns (items, fail0, mids,
    edges) with items the candidate indices that own chunk edges (in edge
    order), fail0 the failure flags before any verdict, mids the smoothing
    candidates and edges the list of (item, from, to) chunks."""
    N, D = path.shape
    n_items = N if smooth else N * 16
    fail = np.zeros(n_items, np.uint32)
    mids = {}
    edges = []
    off = 0
    r = np.float32(rng)
    for item in range(n_items):
        parts = []
        if smooth:
            i = item
            live = False
            if (i & 1) and i < N - 1:
                p0, p1, p2 = path[i - 1], path[i], path[i + 1]
                two = np.float32(2.0)
                mid = (((p0 + p1) / two) + ((p1 + p2) / two)) / two
                mid = mid.astype(np.float32)
                mids[i] = mid
                if sqrt32(sqdist(p1[None], mid)[0]) >= np.float32(min_change):
                    live = True
                    c1 = int(math.ceil(np.float32(l2(p0, mid)) / r))
                    c2 = int(math.ceil(np.float32(l2(p2, mid)) / r))
                    parts = [(p0, mid, c1), (p2, mid, c2)]
            if not live:
                fail[i] = 1
        else:
            i, b = item >> 4, item & 15
            j = i + b + 1
            live = b >= 1 and j < N
            if live:
                d = l2(path[i], path[j])
                c1 = max(int(math.ceil(np.float32(np.float32(d) - np.float32(0.001)) / r)), 0)
                parts = [(path[i], path[j], c1)]
        cnt = sum(p[2] for p in parts)
        if live and cnt > 0:
            if off + cnt > cap:
                fail[item] = 1
            else:
                for frm, to, nc in parts:
                    f, t = chunks(frm, to, nc)
                    edges += [(item, f[c], t[c]) for c in range(nc)]
        off += cnt
    return fail, mids, edges


def apply_smooth(path, fail, mids):
    out = path.copy()
    for i, mid in mids.items():
        if not fail[i]:
            out[i] = mid
    return out


def shortcut(path, fail, best):
    """Shortcut DP and rewrite; returns (path, cost, best)."""
    N = len(path)
    cost = np.full(N, FLT_MAX, np.float32)
    par = np.full(N, -1, np.int64)
    cost[0] = 0
    for i in range(N - 1):
        for b in range(16):
            j = i + b + 1
            if j >= N:
                break
            if b > 0 and fail[i * 16 + b]:
                continue
            c = np.float32(l2(path[i], path[j]) + cost[i])
            if cost[j] > c:
                cost[j] = c
                par[j] = i
    out = path.copy()
    if cost[N - 1] < best:
        child, parent = N - 1, par[N - 1]
        while parent != -1:
            for idx in range(parent + 1, child):
                out[idx] = out[parent]
            child, parent = parent, par[parent]
    return out, cost[N - 1], min(np.float32(best), cost[N - 1])

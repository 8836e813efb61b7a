"""Panda sphere model and base layouts, read from the library's geometry
header (mr.pop_amd/csrc/panda_geometry.h), which is the single source of the
numbers for the HIP kernels, the host code and this package."""
from __future__ import annotations

import re
from functools import lru_cache
from pathlib import Path

import numpy as np

HEADER = Path(__file__).resolve().parent.parent / "csrc" / "panda_geometry.h"


@lru_cache(maxsize=None)
def _tables() -> dict:
    text = HEADER.read_text()
    out = {}
    for m in re.finditer(r"inline constexpr (float|int) (k\w+)(\[[^=]*\])\s*=\s*\{(.*?)\};", text, re.S):
        kind, name, dims, body = m.groups()
        vals = [float(x.rstrip("f")) for x in re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?", body)]
        arr = np.array(vals, np.float32 if kind == "float" else np.int32)
        shape = [int(d) if d.isdigit() else -1 for d in re.findall(r"\[([^\]]*)\]", dims)]
        shape = [s if s != -1 else None for s in shape]
        if None in shape or np.prod(shape) != arr.size:
            arr = arr.reshape(-1, shape[-1]) if shape[-1] and arr.size % shape[-1] == 0 else arr
        else:
            arr = arr.reshape(shape)
        out[name] = arr
    return out


def table(name: str) -> np.ndarray:
    return _tables()[name].copy()


def default_bases(n_arms: int) -> np.ndarray:
    """World <- base poses (n_arms, 12) for 2..5 arms."""
    if n_arms not in (2, 3, 4, 5):
        raise ValueError("default base layouts exist for 2..5 arms")
    return table(f"kBase{n_arms}").reshape(n_arms, 12)

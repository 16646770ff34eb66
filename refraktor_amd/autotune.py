"""Launch-shape autotuner: walkers per block, atomic width, LDS bin-tile size
and walker count, measured on the device through rfk_autotune.

Results are cached per (flame structure, image size, precision) in
`$RFK_CACHE_DIR/tune.json` (default ~/.cache/refraktor_amd) so a batch or a
render loop tunes once.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import _lib
from . import render
from ._lib import check


@dataclass(frozen=True)
class Shape:
    threads_per_block: int = 256
    atomic_width: int = 32
    bin_tile_log2: int = 0
    walkers: int = 1 << 20
    bin_queue: int = 0


# atomic binning with the LDS bin cache at a few shapes, and queue binning
DEFAULT_GRID = ([Shape(tpb, aw, bt, w) for tpb in (256, 1024) for aw in (32, 64) for bt in (11, 12)
                 for w in (1 << 20,)] +
                [Shape(tpb, 64, 0, w, 1) for tpb in (256, 512, 1024) for w in (1 << 19, 1 << 20, 1 << 21)])


def _cache_path() -> Path:
    root = os.environ.get("RFK_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "refraktor_amd")
    return Path(root) / "tune.json"


def _key(flame: render.Flame, width: int, height: int, precision: int) -> str:
    src = flame.source(precision)
    return hashlib.sha256(f"{src}|{width}x{height}|{precision}".encode()).hexdigest()[:32]


def measure(db: render.VariationDB, flame: render.Flame, samples: np.ndarray, width: int, height: int,
            grid: list[Shape] | None = None, probe_iters: int = 1 << 27, precision: int = 0) -> list[dict]:
    """iterations/s of every shape in `grid` (device time); failures keep their status"""
    grid = list(grid or DEFAULT_GRID)
    arr = (_lib.TuneCandidate * len(grid))()
    for i, s in enumerate(grid):
        arr[i].threads_per_block, arr[i].atomic_width = s.threads_per_block, s.atomic_width
        arr[i].bin_tile_log2, arr[i].walkers = s.bin_tile_log2, s.walkers
        arr[i].bin_queue = s.bin_queue
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    best = C.c_size_t()
    check(_lib.lib().rfk_autotune(db.handle, flame.handle, precision, samples.ctypes.data_as(C.POINTER(C.c_double)),
                                  samples.size, width, height, probe_iters, C.cast(arr, C.c_void_p), len(grid),
                                  C.byref(best)))
    out = []
    for i, s in enumerate(grid):
        out.append({**asdict(s), "iters_per_s": arr[i].iters_per_s, "compile_ms": arr[i].compile_ms,
                    "status": arr[i].status, "best": i == best.value})
    return out


def tune(db: render.VariationDB, flame: render.Flame, samples: np.ndarray, width: int, height: int,
         grid: list[Shape] | None = None, probe_iters: int = 1 << 27, precision: int = 0,
         use_cache: bool = True) -> tuple[Shape, list[dict]]:
    """fastest shape for this flame structure and image size (cached)"""
    key = _key(flame, width, height, precision)
    path = _cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        return Shape(**cache[key]["shape"]), cache[key]["results"]
    results = measure(db, flame, samples, width, height, grid, probe_iters, precision)
    best = next(r for r in results if r["best"])
    shape = Shape(best["threads_per_block"], best["atomic_width"], best["bin_tile_log2"], best["walkers"],
                  best["bin_queue"])
    if use_cache:
        cache[key] = {"shape": asdict(shape), "results": results}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            tmp = path.with_suffix(".tmp")
            tmp.write_text(json.dumps(cache))
            tmp.replace(path)
        except OSError:
            pass
    return shape, results

"""Animations: keyframes aligned into one structure, packed as Catmull-Rom
control samples and rendered as a batch in one persistent launch.

The reference renders an animation one frame at a time, interpolating aligned
keyframes by weight (src/refraktor/src/refraktor.cpp:146-198) and packing four
samples per frame for warmup (src/refraktor-gui/src/gui/modals/render_modal.cpp:262-364).
Here every keyframe is packed once; frame f of F covers the segment time
[f, f+1) * nseg / F and is rendered by its own blocks in the batch.
"""
from __future__ import annotations

import math

import numpy as np

from . import autotune
from . import flame as fl
from . import render


def keyframe_flames(n: int, seed: int) -> list[dict]:
    """n seeded keyframes of 6-10 xforms each"""
    return [fl.flame_keyframe(seed * 1000 + i) for i in range(n)]


def control_samples(db: render.VariationDB, keyframes: list, width: int, height: int):
    """aligned flames and the packed control samples [K0, K0, K1, ..., Kn-1, Kn-1]
    (end points doubled so the curve passes through every keyframe)"""
    aligned = db.align(keyframes)
    packed = [f.pack_samples(width, height, 0.0, 0.0, 1) for f in aligned]
    seq = [packed[0]] + packed + [packed[-1]]
    return aligned, np.concatenate(seq)


def frame_budgets(frames: int, seed: int, lo: int = 1 << 26, hi: int = 1 << 29) -> np.ndarray:
    """per-frame iteration budgets, log-uniform in [lo, hi]"""
    rng = fl.SplitMix64(seed)
    a, b = math.log(lo), math.log(hi)
    return np.array([int(math.exp(a + (b - a) * rng.uniform())) for _ in range(frames)], dtype=np.uint64)


class Animation:
    """compile once, render every frame of the keyframe sequence in one batch"""

    def __init__(self, db: render.VariationDB, keyframes: list, width: int, height: int, frames: int, iters,
                 walkers: int = 1 << 18, seed: int = 1, fuse: int = 20, threads_per_block: int = 256,
                 atomic_width: int = 32, bin_tile_log2: int = 0, bin_queue: bool = False):
        self.aligned, self.samples = control_samples(db, keyframes, width, height)
        self.kernel = db.compile(self.aligned[0], threads_per_block=threads_per_block, atomic_width=atomic_width,
                                 bin_tile_log2=bin_tile_log2, bin_queue=bin_queue)
        self.iters = np.ascontiguousarray(np.broadcast_to(np.asarray(iters, dtype=np.uint64), (frames,)))
        self.batch = self.kernel.batch(self.samples, frames, self.iters, width, height, walkers, seed, fuse)

    def render(self, graph: bool = True) -> float:
        """all frames; returns the device time in ms"""
        return self.batch.run(graph)


def batch_grid(walkers: int) -> list:
    """launch shapes a batch can take: queue binning at 256, 512 and 1024
    walkers per workgroup, and atomic binning behind the LDS bin cache"""
    return ([autotune.Shape(tpb, 64, 0, walkers, 1) for tpb in (256, 512, 1024)] +
            [autotune.Shape(1024, 64, 12, walkers, 0)])


def tune_shape(db: render.VariationDB, keyframes: list, width: int, height: int, walkers: int,
               probe_iters: int = 1 << 26, use_cache: bool = True):
    """the fastest batch shape for these keyframes, measured by the
    autotuner on one frame of the aligned structure (its first segment) at
    the batch's size and walkers per frame"""
    aligned, samples = control_samples(db, keyframes, width, height)
    ss = samples.size // (len(keyframes) + 2)
    return autotune.tune(db, aligned[0], samples[:4 * ss], width, height, batch_grid(walkers), probe_iters,
                         use_cache=use_cache)


def main(argv=None) -> int:
    """The animation workload: 4 keyframes of 6-10 xforms, 120 frames at
    1280x720, budgets log-uniform in [2^26, 2^29], queue-binned chunk
    launches (or one persistent launch with --binning atomic) replayed from a
    hipGraph. Prints one JSON line."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--keyframes", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=1 << 18, help="walkers per frame")
    ap.add_argument("--binning", choices=("queue", "atomic"), default="queue")
    ap.add_argument("--tpb", type=int, default=None, help="walkers per workgroup (default 512 queue, 1024 atomic)")
    ap.add_argument("--tile", type=int, default=None,
                    help="LDS bin cache slots (log2) of atomic binning, 0 = off (default 12)")
    ap.add_argument("--atomic-width", type=int, default=64)
    ap.add_argument("--replays", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --binning / --tpb / --tile instead of the autotuned shape")
    args = ap.parse_args(argv)
    db = render.VariationDB()
    keyframes = keyframe_flames(args.keyframes, args.seed)
    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        shape, results = tune_shape(db, keyframes, args.width, args.height, args.walkers, use_cache=False)
        tuned = {"shape": {"binning": "queue" if shape.bin_queue else "atomic",
                           "threads_per_block": shape.threads_per_block, "atomic_width": shape.atomic_width,
                           "bin_tile_log2": shape.bin_tile_log2, "walkers": shape.walkers},
                 "candidates": [{"binning": "queue" if r["bin_queue"] else "atomic",
                                 "threads_per_block": r["threads_per_block"], "bin_tile_log2": r["bin_tile_log2"],
                                 "iters_per_s": r["iters_per_s"], "status": r["status"]} for r in results],
                 "tune_s": time.perf_counter() - t0}
        args.binning = "queue" if shape.bin_queue else "atomic"
        args.tpb, args.tile, args.atomic_width = shape.threads_per_block, shape.bin_tile_log2, shape.atomic_width
    queue = args.binning == "queue"
    if args.tpb is None:
        args.tpb = 512 if queue else 1024
    if args.tile is None:
        args.tile = 0 if queue else 12
    budgets = frame_budgets(args.frames, args.seed)
    t0 = time.perf_counter()
    anim = Animation(db, keyframes, args.width, args.height, args.frames, budgets,
                     walkers=args.walkers, seed=args.seed, threads_per_block=args.tpb,
                     atomic_width=args.atomic_width, bin_tile_log2=args.tile, bin_queue=queue)
    setup_s = time.perf_counter() - t0
    eager_ms = anim.render(graph=False)
    graph_ms = [anim.render(graph=True) for _ in range(args.replays)]
    passes = sum(anim.batch.frame(f)["total_passes"] for f in range(args.frames))
    best = min(graph_ms)
    print(json.dumps({"workload": f"{args.keyframes} keyframes, {args.frames} frames at {args.width}x{args.height}, "
                                  "budgets log-uniform in [2^26, 2^29], " +
                                  ("queue-binned chunks" if queue else "one persistent launch") + ", hipGraph replay",
                      "iterations": int(budgets.sum()), "passes": int(passes), "setup_s": setup_s,
                      "eager_ms": eager_ms, "graph_ms": graph_ms, "frames_per_s": args.frames / (best * 1e-3),
                      "iterations_per_s": passes / (best * 1e-3),
                      "config": {"binning": args.binning, "walkers": args.walkers, "threads_per_block": args.tpb,
                                 "bin_tile_log2": args.tile, "grid": anim.batch.grid, "items": anim.batch.items,
                                 "atomic_width": args.atomic_width},
                      "autotune": tuned}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

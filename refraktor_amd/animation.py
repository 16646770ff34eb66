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
                 atomic_width: int = 32):
        self.aligned, self.samples = control_samples(db, keyframes, width, height)
        self.kernel = db.compile(self.aligned[0], threads_per_block=threads_per_block, atomic_width=atomic_width)
        self.iters = np.ascontiguousarray(np.broadcast_to(np.asarray(iters, dtype=np.uint64), (frames,)))
        self.batch = self.kernel.batch(self.samples, frames, self.iters, width, height, walkers, seed, fuse)

    def render(self, graph: bool = True) -> float:
        """all frames; returns the device time in ms"""
        return self.batch.run(graph)

"""Train the U-Net denoiser on pairs the renderer produces.

    python -m refraktor_amd.train_denoiser --flames 256 --steps 4000 --out weights.bin

The shipped weights (weights/unet_1080.bin) came from
    python -m refraktor_amd.train_denoiser --flames 600 --holdout 8 --steps 60000 \
        --batch 16 --lr 2e-3 --low 33554432 --de-radius 0 --out unet_1080.bin

Pairs: random seeded flames (keyframe flames of 6-10 xforms and 1080p flames
with a final xform, alternating), the low-budget render as input features and
the converged render as target, both through density estimation and the
log-density tonemap. Flame kernels are compiled by a thread pool (hiprtc runs
on host cores); everything else stays on the device.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
import multiprocessing as mp
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

from . import _lib
from . import denoise
from . import flame as fl
from . import render
from ._lib import check

TRAIN_SEED0 = 100000  # training flames never share seeds with the benchmark flames


def training_flame(seed: int) -> dict:
    return fl.flame_keyframe(seed) if seed % 2 == 0 else fl.flame_1080(seed)


# launch shape of the renders: 1024 walkers per block, queue binning, 64-bit
# bins so the converged targets cannot wrap
SHAPE = dict(threads_per_block=1024, atomic_width=64, bin_queue=True, hist_bits=64)


def _compile_only(flame: dict) -> int:
    """hiprtc compile into the code-object cache; runs in a worker process
    that never touches the GPU"""
    db = render.VariationDB()
    s = {k: v for k, v in SHAPE.items() if k != "threads_per_block"}
    return db.flame(flame).compile_check(0, SHAPE["threads_per_block"], **s)


def compile_all(db: render.VariationDB, flames: list[dict], threads: int = 16, log=print) -> list:
    """hiprtc in `threads` worker processes (it does not compile in parallel
    inside one process), then load every kernel from the cache here"""
    t0 = time.time()
    with ProcessPoolExecutor(max_workers=threads, mp_context=mp.get_context("spawn")) as ex:
        for i, _ in enumerate(ex.map(_compile_only, flames, chunksize=4)):
            if i % 32 == 31:
                log(f"[pairs] compiled {i + 1}/{len(flames)} in {time.time() - t0:.1f} s")
    return [db.compile(f, **SHAPE) for f in flames]


def coarse_importance(rgb: np.ndarray) -> np.ndarray:
    """16x16-cell weights: mean brightness plus local contrast, floored"""
    h, w = rgb.shape[0] // 16 * 16, rgb.shape[1] // 16 * 16
    cells = rgb[:h, :w].mean(axis=-1).reshape(h // 16, 16, w // 16, 16)
    m = cells.mean(axis=(1, 3))
    s = cells.std(axis=(1, 3))
    return (m + 2.0 * s + 0.02).astype(np.float64)


def make_pairs(db, seeds, width, height, low_iters, high_iters, de_radius, threads=16, log=print):
    flames = [training_flame(s) for s in seeds]
    t0 = time.time()
    kernels = compile_all(db, flames, threads, log)
    log(f"[pairs] compiled {len(kernels)} flame kernels in {time.time() - t0:.1f} s")
    n, npx = len(seeds), width * height
    feat = render.DeviceBuffer(n * npx * denoise.FEAT_C * 2)
    base = render.DeviceBuffer(n * npx * 12)
    tgt = render.DeviceBuffer(n * npx * 12)
    lib = _lib.lib()
    importance = []
    t0 = time.time()
    for i, (s, f, k) in enumerate(zip(seeds, flames, kernels)):
        lo = denoise.render_frame(db, f, width, height, low_iters, seed=s * 7 + 1, de_radius=de_radius, kernel=k)
        hi = denoise.render_frame(db, f, width, height, high_iters, seed=s * 7 + 2, de_radius=de_radius, kernel=k)
        fe, ba = denoise.features(lo)
        t = denoise.rgb_of(hi)
        check(lib.rfk_memcpy_dtod(feat.ptr + i * npx * denoise.FEAT_C * 2, fe.ptr, npx * denoise.FEAT_C * 2))
        check(lib.rfk_memcpy_dtod(base.ptr + i * npx * 12, ba.ptr, npx * 12))
        check(lib.rfk_memcpy_dtod(tgt.ptr + i * npx * 12, t.ptr, npx * 12))
        importance.append(coarse_importance(t.download(np.float32, (height, width, 3))))
        if i % 16 == 0 or i == n - 1:
            log(f"[pairs] {i + 1}/{n} rendered, {time.time() - t0:.1f} s")
    return denoise.Dataset(feat, base, tgt, n, height, width), np.stack(importance)


def evaluate(den: denoise.Denoiser, data: denoise.Dataset, idx: list[int]) -> dict:
    """PSNR of the low-budget image and of the denoised image vs the target"""
    npx = data.width * data.height
    out = render.DeviceBuffer(npx * 12)
    lib = _lib.lib()
    res = {"psnr_low": [], "psnr_denoised": []}
    for i in idx:
        fe = render.DeviceBuffer(npx * denoise.FEAT_C * 2)
        ba = render.DeviceBuffer(npx * 12)
        check(lib.rfk_memcpy_dtod(fe.ptr, data.feat.ptr + i * npx * denoise.FEAT_C * 2, npx * denoise.FEAT_C * 2))
        check(lib.rfk_memcpy_dtod(ba.ptr, data.base.ptr + i * npx * 12, npx * 12))
        den.forward(fe, ba, 1, data.height, data.width, out)
        got = out.download(np.float32, (data.height, data.width, 3))
        low = ba.download(np.float32, (data.height, data.width, 3))
        tgt = render.download(data.target.ptr + i * npx * 12, np.float32, (data.height, data.width, 3))
        res["psnr_low"].append(denoise.psnr(low, tgt))
        res["psnr_denoised"].append(denoise.psnr(got, tgt))
    return {k: float(np.mean(v)) for k, v in res.items()} | {"frames": len(idx)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--flames", type=int, default=256)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--low", type=int, default=1 << 25)
    ap.add_argument("--high", type=int, default=1 << 32)
    ap.add_argument("--de-radius", type=float, default=0.0)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--init", type=str, default=None, help="start from these weights")
    ap.add_argument("--out", type=str, required=True)
    ap.add_argument("--threads", type=int, default=16)
    args = ap.parse_args(argv)

    db = render.VariationDB()
    seeds = list(range(TRAIN_SEED0, TRAIN_SEED0 + args.flames + args.holdout))
    data, importance = make_pairs(db, seeds, args.width, args.height, args.low, args.high, args.de_radius,
                                  args.threads)
    ntrain = args.flames
    train_view = denoise.Dataset(data.feat, data.base, data.target, ntrain, data.height, data.width)
    den = denoise.Denoiser(seed=1, weights=args.init)
    before = evaluate(den, data, list(range(ntrain, ntrain + args.holdout))) if args.init else None
    t0 = time.time()
    losses = denoise.train(den, train_view, args.steps, batch=args.batch, crop=args.crop, lr=args.lr,
                           importance=importance[:ntrain])
    train_s = time.time() - t0
    den.save(args.out)
    after = evaluate(den, data, list(range(ntrain, ntrain + args.holdout)))
    print(json.dumps({"train_seconds": train_s, "steps": args.steps, "final_loss": float(np.mean(losses[-100:])),
                      "holdout_before": before, "holdout_after": after, "out": args.out,
                      "flames": args.flames, "low": args.low, "high": args.high}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark: 1080p plain fractal-flame frames on one MI355X.

One step = one full frame of the 1920x1080 workload (8 xforms of 2-4
variations + final xform, seeded synthetic flame): warmup (sample
evaluation, precalc, walker seeding, 20 fuse passes), 2^30 binned iterations,
density estimation (radius 9) and a log-density tonemap to half4.

Prints one JSON line: frames/s (value), iterations/s, PSNR of the frame vs a
converged 2^32-iteration render, the iterate kernel's roofline and the host
baseline (the oracle's OpenMP render of the same flame, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def psnr(a, b) -> float:
    import numpy as np
    mse = float(np.mean((a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(1.0 / mse)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tpb", type=int, default=256)
    ap.add_argument("--walkers", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baseline")
    ap.add_argument("--no-psnr", action="store_true", help="skip the converged reference render")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np

    from refraktor_amd import _lib
    from refraktor_amd import flame as fl
    from refraktor_amd import render

    W, H = 1920, 1080
    ITERS = 1 << 30
    DE_R = 9.0
    f = fl.flame_1080(args.seed)
    db = render.VariationDB()
    k = db.compile(f, threads_per_block=args.tpb)
    samples = k.pack_samples(W, H)
    gamma, bright, vib = f["gamma"], f["brightness"], f["vibrancy"]

    def frame(seed: int):
        st = k.warmup(samples, W, H, seed=seed, fuse=20, walkers=args.walkers)
        res = st.bin(iters=ITERS)
        de = st.density_estimate(DE_R, 0.0, 0.4)
        img = render.tonemap_image(de, gamma, bright, vib, st.info().quality, fmt=_lib.PIXEL_HALF4)
        return st, res, img

    for i in range(args.warmup):
        frame(args.seed + 1000 + i)
    _lib.lib().rfk_synchronize()
    t0 = time.perf_counter()
    kernel_ms, draws, passes = [], [], []
    last = None
    for i in range(args.steps):
        st, res, img = frame(args.seed + i)
        kernel_ms.append(res["kernel_ms"])
        draws.append(res["total_draws"])
        passes.append(res["total_passes"])
        last = (st, res, img)
    _lib.lib().rfk_synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    fps = 1e3 / ms_per_step

    # iterate kernel: algorithmic bytes = 16 B per draw (one RGBA u32 increment)
    kms = float(np.mean(kernel_ms))
    dr = float(np.mean(draws))
    achieved_gbs = dr * 16.0 / (kms * 1e-3) / 1e9
    it_per_s = float(np.mean(passes)) / (kms * 1e-3)
    # measured atomic peak at the same footprint (RGBA u32, 1080p)
    peak_inc = 0.0
    try:
        import ctypes as C
        v = C.c_double()
        _lib.check(_lib.lib().rfk_atomic_peak(W * H * 16, 1 << 28, C.byref(v)))
        peak_inc = v.value
    except Exception:
        pass

    psnr_db = None
    if not args.no_psnr:
        st_ref = k.warmup(samples, W, H, seed=args.seed + 5000, fuse=20, walkers=args.walkers)
        st_ref.bin(iters=1 << 32)
        de = st_ref.density_estimate(DE_R, 0.0, 0.4)
        ref = render.tonemap_image(de, gamma, bright, vib, st_ref.info().quality, fmt=_lib.PIXEL_FLOAT4)
        psnr_db = psnr(last[2].astype(np.float32), ref)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(f, W, H, args.seed, it_per_frame=ITERS)

    line = {
        "metric": "1080p frames/sec at PSNR>=40dB vs converged plain render; iterations/sec",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded flame (SplitMix64 seed %d)" % args.seed,
        "config": {"workload": "plain render 1920x1080, 8 xforms x 2-4 variations + final xform, 2^30 iterations, "
                               "DE radius 9, log-density tonemap, vibrancy 1",
                   "width": W, "height": H, "iterations": ITERS, "walkers": last[0].info().walkers,
                   "threads_per_block": args.tpb, "path": "plain"},
        "iterations_per_s": it_per_s,
        "iterate_kernel_ms": kms,
        "psnr_db_vs_converged": psnr_db,
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "rfk_iterate", "bytes_per_draw": 16, "draws_per_launch": dr,
                     "atomic_peak_increments_per_s": peak_inc,
                     "atomic_frac": (dr * 4.0 / (kms * 1e-3) / peak_inc) if peak_inc else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(f: dict, W: int, H: int, seed: int, it_per_frame: int) -> dict:
    """the oracle's OpenMP render of the same flame on a bounded sample"""
    from oracle import oracle
    from refraktor_amd import render
    blocks, tpb = 1024, 256
    passes = 256  # 1024*256*256 = 2^26 iterations
    L, P32, pals = oracle.prepare(f, render.CONFIG_DIR, W, H, blocks, seed)
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    _, draws = oracle.render(L, P32, P32.shape[1], pals, blocks, tpb, W, H, seed, 20, passes, threads=threads)
    dt = time.perf_counter() - t0
    iters = blocks * tpb * passes
    ips = iters / dt
    return {"value": ips / it_per_frame, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{iters} iterations (+20 fuse passes) of the same 1080p flame on {threads} threads in "
                      f"{dt:.2f} s ({ips:.3e} it/s), scaled to a 2^30-iteration frame",
            "iterations_per_s": ips}


if __name__ == "__main__":
    sys.exit(main())

"""Time the GPU JPEG encoder on a synthetic frame (the JPEG stage behind
rfkt::jpeg_encoder, interface/jpeg_encoder.h:11-24). Prints one JSON line.

    python -m refraktor_amd.jpeg [--width 1920] [--height 1080] [--quality 90]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, render


def synthetic_frame(width: int, height: int, seed: int = 1) -> np.ndarray:
    """gradients with a noise patch: smooth blocks and blocks of full 10-bit AC"""
    y, x = np.mgrid[0:height, 0:width]
    img = np.stack([x * 255 // width, y * 255 // height, (x ^ y) & 255], -1).astype(np.uint8)
    ph, pw = height * 5 // 18, width * 5 // 24
    img[:ph, :pw] = np.random.default_rng(seed).integers(0, 256, size=(ph, pw, 3), dtype=np.uint8)
    return img


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--quality", type=int, nargs="+", default=[90, 75])
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    img = synthetic_frame(args.width, args.height)
    buf = render.DeviceBuffer(img.nbytes)
    buf.upload(img)
    enc = render.JpegEncoder()
    out = {"workload": f"{args.width}x{args.height} RGB8 device image -> baseline JFIF (4:4:4)", "results": []}
    for q in args.quality:
        ts = []
        for _ in range(args.reps):
            t = time.perf_counter()
            data = enc.encode_device(buf.ptr, args.width, args.height, q, _lib.PIXEL_RGB8)
            ts.append(time.perf_counter() - t)
        ts.sort()
        out["results"].append({"quality": q, "bytes": len(data), "best_ms": ts[0] * 1e3,
                               "median_ms": ts[len(ts) // 2] * 1e3})
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

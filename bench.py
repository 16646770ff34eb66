"""Benchmark: 1080p fractal-flame frames on one MI355X.

Workload: the 1920x1080 flame (8 xforms of 2-4 variations + final xform,
seeded synthetic flame). Two paths, each timed over --steps frames:
  plain     warmup (sample evaluation, precalc, walker seeding, 20 fuse
            passes), 2^30 binned iterations, density estimation (radius 9),
            log-density tonemap to half4;
  denoised  2^25 iterations, the log-density tonemap of the histogram itself
            (no density estimation: the network, trained on such frames,
            does the smoothing), then the U-Net (features, forward) to RGB.
Both are scored by PSNR against a converged 2^32-iteration render of the same
flame. `value` is the frames/s of the denoised path when it reaches 40 dB
(else of the plain path), the metric "1080p frames/sec at PSNR>=40dB".

Frames are independent renders (seed per frame); --pipeline N keeps N frames
in flight, each on its own streams and buffers, for both paths alike, so one
frame's U-Net and density estimation run beside the next frame's binning.
Each path is also timed one frame at a time (`sequential_ms`, the latency).

Prints one JSON line with both paths, iterations/s, the binning roofline
(16 B of histogram increment per draw over the bin call's device time; the
HBM traffic of the same call from the committed rocprofv3 counter summary),
the denoiser's MFMA roofline (the U-Net forward's FLOPs over its device time)
and the host baseline (the oracle's OpenMP render of the same flame on a
bounded sample).
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


def check(rc: int) -> None:
    from refraktor_amd import _lib
    _lib.check(rc)


def psnr(a, b) -> float:
    import numpy as np
    mse = float(np.mean((a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(1.0 / mse)


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def unet_flops(h: int, w: int) -> float:
    """multiply-adds x 2 of the denoiser's 15 3x3 convolutions at h x w
    (widths 32/64/96/128, 8 input and 8 output channels, skips concatenated)"""
    c = (32, 64, 96, 128)
    lv = [(h >> i, w >> i) for i in range(4)]
    layers = [(0, 8, c[0]), (0, c[0], c[0]), (1, c[0], c[1]), (1, c[1], c[1]), (2, c[1], c[2]), (2, c[2], c[2]),
              (3, c[2], c[3]), (3, c[3], c[3]), (2, c[3] + c[2], c[2]), (2, c[2], c[2]), (1, c[2] + c[1], c[1]),
              (1, c[1], c[1]), (0, c[1] + c[0], c[0]), (0, c[0], c[0]), (0, c[0], 8)]
    return float(sum(2 * 9 * lv[l][0] * lv[l][1] * ci * co for l, ci, co in layers))


def unet_roofline(lane, h: int, w: int, reps: int = 10) -> dict:
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    ev0, ev1 = C.c_void_p(), C.c_void_p()

    def check_hip(err: int) -> None:
        if err != 0:
            raise RuntimeError(f"HIP error {err} while timing the U-Net forward")

    check_hip(hip.hipEventCreate(C.byref(ev0)))
    check_hip(hip.hipEventCreate(C.byref(ev1)))
    lane.den.forward(lane.feat, lane.base, 1, h, w, lane.rgb)  # warm
    check_hip(hip.hipEventRecord(ev0, None))
    for _ in range(reps):
        lane.den.forward(lane.feat, lane.base, 1, h, w, lane.rgb)
    check_hip(hip.hipEventRecord(ev1, None))
    check_hip(hip.hipEventSynchronize(ev1))
    ms = C.c_float()
    check_hip(hip.hipEventElapsedTime(C.byref(ms), ev0, ev1))
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    per = ms.value / reps
    flops = unet_flops(h, w)
    tflops = flops / (per * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
            "kernel": "U-Net forward: 15 conv3x3 (BF16 MFMA, FP32 accumulate) + 3 average pools",
            "flops_per_launch": flops, "launch_ms": per}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tpb", type=int, default=512, help="plain path walkers per workgroup (without autotuning)")
    ap.add_argument("--den-tpb", type=int, default=512, help="denoised path walkers per workgroup (without autotuning)")
    ap.add_argument("--den-binning", choices=("queue", "atomic"), default="queue",
                    help="denoised path binning (without autotuning)")
    ap.add_argument("--den-tile", type=int, default=12, help="denoised path LDS bin cache slots (log2), atomic binning")
    ap.add_argument("--no-autotune", action="store_true",
                    help="plain path at --tpb / --binning / --walkers instead of the autotuned shape")
    ap.add_argument("--binning", choices=("queue", "atomic"), default="queue",
                    help="queue: records sorted by screen tile and summed in LDS; atomic: LDS bin cache + atomics")
    ap.add_argument("--tile", type=int, default=12, help="LDS bin cache slots (log2) of atomic binning, 0 = off")
    ap.add_argument("--atomic-width", type=int, default=64)
    ap.add_argument("--hist-bits", type=int, default=32)
    ap.add_argument("--walkers", type=int, default=1 << 20, help="walkers of the plain path")
    ap.add_argument("--low-walkers", type=int, default=1 << 19,
                    help="walkers of the denoised path")
    ap.add_argument("--low-iters", type=int, default=1 << 25, help="budget of the denoised path")
    ap.add_argument("--low-fuse", type=int, default=20, help="fuse iterations per walker of the denoised path")
    ap.add_argument("--low-de-radius", type=float, default=0.0,
                    help="density estimation radius of the denoised path (0: tonemap the histogram itself)")
    ap.add_argument("--weights", type=str, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baseline")
    ap.add_argument("--no-psnr", action="store_true", help="skip the converged reference render")
    ap.add_argument("--no-denoise", action="store_true", help="plain path only")
    ap.add_argument("--pipeline", type=int, default=3,
                    help="frames in flight, each on its own streams (both paths alike); 1 = one frame at a time")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only three 1 GiB device-to-device copies: the known-size traffic that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of the bin call into bytes")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure the bin call's HBM traffic")
    ap.add_argument("--unet-only", action="store_true",
                    help="time the U-Net forward alone (--steps forwards of one denoised frame's features)")
    ap.add_argument("--tune-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2
    if args.copy_calibration:
        from refraktor_amd import _lib, render
        n = 1 << 30
        src, dst = render.DeviceBuffer(n), render.DeviceBuffer(n)
        src.zero()
        for _ in range(3):
            _lib.check(_lib.lib().rfk_memcpy_dtod(dst.ptr, src.ptr, n))
        _lib.check(_lib.lib().rfk_synchronize())
        print(json.dumps({"copy_calibration": {"bytes_per_copy": n, "copies": 3}}))
        return 0

    # The bin call's launch shape (autotuned) and its HBM traffic (rocprofv3
    # counters) come from child processes started before this process
    # touches the GPU: the counter runs need the profiler around a process
    # of their own.
    tuned = traffic = None
    if not (args.tune_only or args.traffic_probe or args.unet_only):
        if not args.no_autotune:
            child = run_child(["--tune-only", "--seed", str(args.seed), "--low-walkers", str(args.low_walkers),
                               "--low-iters", str(args.low_iters)])
            tuned = child["autotune"]
            tuned["denoised"] = child["autotune_denoised"]
            dsh = tuned["denoised"]["shape"]
            args.den_binning, args.den_tpb, args.den_tile = dsh["binning"], dsh["threads_per_block"], dsh["bin_tile_log2"]
            sh = tuned["shape"]
            args.binning, args.tpb, args.tile = sh["binning"], sh["threads_per_block"], sh["bin_tile_log2"]
            args.atomic_width, args.walkers = sh["atomic_width"], sh["walkers"]
            args.no_autotune = True
        traffic = {"note": "not measured (--no-traffic)"} if args.no_traffic else live_traffic(args)

    import ctypes as C

    import numpy as np

    from refraktor_amd import _lib
    from refraktor_amd import denoise
    from refraktor_amd import flame as fl
    from refraktor_amd import render

    W, H = 1920, 1080
    ITERS = 1 << 30
    DE_R = 9.0
    f = fl.flame_1080(args.seed)
    db = render.VariationDB()
    if args.traffic_probe:
        return traffic_probe(args, db, f, W, H, ITERS)
    if not (args.no_autotune or args.unet_only):
        # the plain path's launch shape from the autotuner: one 2^28-iteration
        # bin call of this flame per candidate, device time
        from refraktor_amd import autotune
        grid = ([autotune.Shape(tpb, 64, 0, w, 1) for tpb in (512, 1024) for w in (1 << 20, 1 << 21)] +
                [autotune.Shape(1024, 64, 12, 1 << 20, 0)])
        fo = db.flame(f)
        t0 = time.perf_counter()
        shape, results = autotune.tune(db, fo, fo.pack_samples(W, H), W, H, grid, 1 << 28, use_cache=False)
        args.binning = "queue" if shape.bin_queue else "atomic"
        args.tpb, args.tile, args.atomic_width, args.walkers = (shape.threads_per_block, shape.bin_tile_log2,
                                                                shape.atomic_width, shape.walkers)
        tuned = {"shape": {"binning": args.binning, "threads_per_block": args.tpb, "bin_tile_log2": args.tile,
                           "atomic_width": args.atomic_width, "walkers": args.walkers},
                 "candidates": [{"binning": "queue" if r["bin_queue"] else "atomic",
                                 "threads_per_block": r["threads_per_block"], "bin_tile_log2": r["bin_tile_log2"],
                                 "walkers": r["walkers"], "iters_per_s": r["iters_per_s"]} for r in results],
                 "tune_s": time.perf_counter() - t0}
        if args.tune_only:
            # the denoised path's shape: one bin call of its budget per
            # candidate at its walker count (queue or atomic binning; the
            # histogram is the same either way)
            dgrid = ([autotune.Shape(tpb, 64, 0, args.low_walkers, 1) for tpb in (256, 512, 1024)] +
                     [autotune.Shape(tpb, 64, t, args.low_walkers, 0) for tpb in (512, 1024) for t in (11, 12)])
            t0 = time.perf_counter()
            dshape, dres = autotune.tune(db, fo, fo.pack_samples(W, H), W, H, dgrid, args.low_iters, use_cache=False)
            den = {"shape": {"binning": "queue" if dshape.bin_queue else "atomic",
                             "threads_per_block": dshape.threads_per_block,
                             "bin_tile_log2": 0 if dshape.bin_queue else dshape.bin_tile_log2,
                             "walkers": dshape.walkers},
                   "candidates": [{"binning": "queue" if r["bin_queue"] else "atomic",
                                   "threads_per_block": r["threads_per_block"], "bin_tile_log2": r["bin_tile_log2"],
                                   "walkers": r["walkers"], "iters_per_s": r["iters_per_s"]} for r in dres],
                   "tune_s": time.perf_counter() - t0}
            print(json.dumps({"autotune": tuned, "autotune_denoised": den}), flush=True)
            return 0
    queue = args.binning == "queue"
    if queue:
        args.tile = 0
    k = db.compile(f, threads_per_block=args.tpb, atomic_width=args.atomic_width, bin_tile_log2=args.tile,
                   hist_bits=args.hist_bits, bin_queue=queue)
    k64 = db.compile(f, threads_per_block=args.tpb, bin_tile_log2=args.tile, hist_bits=64, bin_queue=queue)
    # the denoised path's small budget at its own shape
    den_queue = args.den_binning == "queue"
    den_tile = 0 if den_queue else args.den_tile
    kd = db.compile(f, threads_per_block=args.den_tpb, atomic_width=64, bin_tile_log2=den_tile, hist_bits=32,
                    bin_queue=den_queue)
    samples = k.pack_samples(W, H)
    gamma, bright, vib = f["gamma"], f["brightness"], f["vibrancy"]
    lib = _lib.lib()
    npx = W * H

    weights = Path(args.weights) if args.weights else denoise.DEFAULT_WEIGHTS
    use_den = not args.no_denoise and weights.exists()

    class Lane:
        """one frame in flight: its states (re-warmed for every frame), images,
        feature buffers and denoiser; all of a frame's work runs on its
        state's stream, so frames of different lanes overlap on the device"""

        def __init__(self):
            self.states, self.images = {}, {}
            self.de = render.DeviceBuffer(npx * 16)
            self.den = denoise.Denoiser(weights=weights) if use_den else None
            self.feat = render.DeviceBuffer(npx * denoise.FEAT_C * 2)
            self.base = render.DeviceBuffer(npx * 12)
            self.rgb = render.DeviceBuffer(npx * 12)
            self.pending = []  # states whose bin call nobody has waited for yet

    lanes = [Lane() for _ in range(max(1, args.pipeline))]

    def settle(lane: Lane) -> list:
        """wait for the lane's outstanding bin calls (their results)"""
        out = [st.bin_wait() for st in lane.pending]
        lane.pending.clear()
        return out

    def plain(lane: Lane, seed: int, kern=k, iters=ITERS, fmt=_lib.PIXEL_HALF4, walkers=args.walkers, de_r=DE_R,
              smp=samples, tone=(gamma, bright, vib), fuse=20):
        settle(lane)
        st = lane.states.get((id(kern), walkers))
        if st is None:
            st = lane.states[(id(kern), walkers)] = kern.warmup(smp, W, H, seed=seed, fuse=fuse, walkers=walkers)
        else:
            st.rewarm(smp, seed=seed, fuse=fuse)
        # density estimation is enqueued behind the bin call; the host only
        # waits for the call's counters (the tonemap needs its quality)
        st.bin_async(iters=iters)
        if de_r > 0:
            check(lib.rfk_density_estimate(st.handle, de_r, 0.0, 0.4, lane.de.ptr))
        res = st.bin_wait()
        out = lane.images.get(fmt)
        if out is None:
            out = lane.images[fmt] = render.DeviceBuffer(npx * _lib.PIXEL_BYTES[fmt])
        inf = st.info()
        ta = _lib.TonemapArgs(inf.quality, tone[0], tone[1], tone[2], 0.0)
        # everything of a frame on the state's stream: the next re-warm waits for it
        if de_r > 0:
            check(lib.rfk_tonemap_image(lane.de.ptr, W, H, out.ptr, fmt, C.byref(ta), inf.stream))
        else:
            check(lib.rfk_tonemap(st.handle, out.ptr, fmt, C.byref(ta)))  # the histogram itself
        return st, res, out

    den_kernel = {"k": kd}

    def denoised(lane: Lane, seed: int, kern=None, smp=samples, tone=(gamma, bright, vib)):
        kern = kern or den_kernel["k"]
        if args.low_de_radius > 0:
            st, res, tm = plain(lane, seed, kern, args.low_iters, _lib.PIXEL_FLOAT4, args.low_walkers,
                                args.low_de_radius, smp, tone, args.low_fuse)
            inf = st.info()
            check(lib.rfk_denoiser_features(inf.histogram, inf.hist_bits, tm.ptr, npx, lane.feat.ptr, lane.base.ptr,
                                            inf.stream))
            lane.den.forward(lane.feat, lane.base, 1, H, W, lane.rgb, stream=inf.stream)
            return st, res, lane.rgb
        # the whole frame is enqueued without waiting for its bin call: the
        # tonemap takes the call's quality from its counters on the device,
        # and the host collects the call's result when the lane comes round
        # again, so the next frames' binning is queued behind this one's
        # network on their own streams
        settle(lane)
        key = (id(kern), args.low_walkers)
        st = lane.states.get(key)
        if st is None:
            st = lane.states[key] = kern.warmup(smp, W, H, seed=seed, fuse=args.low_fuse, walkers=args.low_walkers)
        else:
            st.rewarm(smp, seed=seed, fuse=args.low_fuse)
        st.bin_async(iters=args.low_iters)
        ta = _lib.TonemapArgs(0.0, tone[0], tone[1], tone[2], 0.0)
        check(lib.rfk_tonemap_features(st.handle, C.byref(ta), lane.feat.ptr, lane.base.ptr))
        lane.den.forward(lane.feat, lane.base, 1, H, W, lane.rgb, stream=st.info().stream)
        lane.pending.append(st)
        return st, None, lane.rgb

    def timed(fn, steps: int, warm: int, nl: int = len(lanes)):
        """`steps` frames round-robin over the first `nl` lanes"""
        for i in range(max(warm, nl)):  # every lane warm
            fn(lanes[i % nl], args.seed + 1000 + i)
        lib.rfk_synchronize()
        t0 = time.perf_counter()
        kms, draws, passes, last = [], [], [], None
        for i in range(steps):
            st, res, img = fn(lanes[i % nl], args.seed + i)
            if res is not None:
                kms.append(res["kernel_ms"])
                draws.append(res["total_draws"])
                passes.append(res["total_passes"])
            last = img
        lib.rfk_synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        for lane in lanes:
            for res in settle(lane):
                kms.append(res["kernel_ms"])
                draws.append(res["total_draws"])
                passes.append(res["total_passes"])
        return ms, float(np.mean(kms)), float(np.mean(draws)), float(np.mean(passes)), last

    if args.unet_only:
        if not use_den:
            raise SystemExit(f"--unet-only: no trained weights at {weights}")
        denoised(lanes[0], args.seed)
        lib.rfk_synchronize()
        print(json.dumps({"roofline_denoiser": unet_roofline(lanes[0], H, W, reps=args.steps)}), flush=True)
        return 0

    if use_den and tuned is not None and "denoised" in tuned:
        # the denoised path's workgroup size by frames in flight: a smaller
        # iterate workgroup fits beside another frame's convolutions on a CU,
        # which the bin call's own rate cannot show
        # (two rounds over the candidates, the best of each candidate's two
        # timings: clocks drift between candidates timed one after another)
        frame_ms = {}
        cands = sorted({c["threads_per_block"] for c in tuned["denoised"]["candidates"] if c["binning"] == "queue"})
        kerns = {tpb: db.compile(f, threads_per_block=tpb, atomic_width=64, hist_bits=32, bin_queue=True)
                 for tpb in cands}
        for _ in range(2):
            for tpb, kk in kerns.items():
                ms = timed(lambda lane, seed, kk=kk: denoised(lane, seed, kk), 30, 5)[0]
                frame_ms[tpb] = min(frame_ms.get(tpb, ms), ms)
                for lane in lanes:  # free the candidate's states (their record buffers)
                    lane.states = {key: st for key, st in lane.states.items() if key[0] != id(kk)}
        best = min(frame_ms, key=frame_ms.get)
        args.den_binning, args.den_tpb, den_tile = "queue", best, 0
        kd = den_kernel["k"] = db.compile(f, threads_per_block=best, atomic_width=64, hist_bits=32, bin_queue=True)
        tuned["denoised"]["frames_in_flight_ms"] = {str(k): v for k, v in frame_ms.items()}
        tuned["denoised"]["shape"].update({"binning": "queue", "threads_per_block": best, "bin_tile_log2": 0})

    ms_plain, kms, dr, ps, img_plain = timed(plain, args.steps, args.warmup)
    ms_den = img_den = None
    if use_den:
        ms_den, _, _, _, img_den = timed(denoised, args.steps, args.warmup)

    psnr_plain = psnr_den = None
    judge = held_out = None
    if not args.no_psnr:
        # the judge: the converged plain render, 2^32 iterations into u64 bins
        # (no wrap in hot pixels), of the same flame. Renders are
        # deterministic (integer bins, counter-based streams), so the timed
        # frames are scored by rendering their seeds again, untimed.
        _, _, ref_buf = plain(lanes[0], args.seed + 5000, k64, 1 << 32, _lib.PIXEL_FLOAT4)
        ref = ref_buf.download(np.float32, (H, W, 4))
        frames_plain, frames_den = [], []
        for i in range(args.steps):
            _, _, img = plain(lanes[0], args.seed + i)
            lib.rfk_synchronize()
            frames_plain.append(score(img.download(np.float16, (H, W, 4)).astype(np.float32), ref))
            if use_den:
                _, _, img = denoised(lanes[0], args.seed + i)
                lib.rfk_synchronize()
                frames_den.append(score(img.download(np.float32, (H, W, 3)), ref))
        judge = {"reference": "plain render of the same flame, 2^32 iterations, DE radius 9",
                 "plain": summarize(frames_plain), "denoised": summarize(frames_den) if use_den else None}
        psnr_plain = judge["plain"]["psnr_db_min"]
        if use_den:
            psnr_den = judge["denoised"]["psnr_db_min"]
            held_out = held_out_flames(args, db, lanes, plain, denoised, score, summarize)

    # one frame at a time: the per-frame latency of each path, and the bin
    # call's device time without other frames' kernels beside it
    if len(lanes) > 1:
        seq_plain, kms, dr, ps, _ = timed(plain, args.steps, args.warmup, 1)
    else:
        seq_plain = ms_plain
    seq_den = None
    if use_den:
        seq_den = timed(denoised, args.steps, args.warmup, 1)[0] if len(lanes) > 1 else ms_den

    # the U-Net forward alone: its FLOPs over its device time (HIP events on
    # one stream, ten forwards of the last denoised frame's features)
    unet = None
    if use_den:
        unet = unet_roofline(lanes[0], H, W)

    # binning: algorithmic bytes = 16 B per draw (one RGBA u32 increment) over
    # the bin call's device time (HIP events on the state's stream: the
    # iterate launches and, with queue binning, the sort and tile-sum kernels)
    achieved_gbs = dr * 16.0 / (kms * 1e-3) / 1e9
    it_per_s = ps / (kms * 1e-3)
    peak_inc = 0.0
    try:
        v = C.c_double()
        check(lib.rfk_atomic_peak(npx * 16, 1 << 28, C.byref(v)))
        peak_inc = v.value
    except Exception:
        pass
    if traffic and traffic.get("bytes_per_bin_call"):
        # the counter run's bytes per draw times this run's draws per call
        traffic_bytes = traffic["bytes_per_bin_call"] / traffic["draws"] * dr
    else:
        traffic_bytes = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(f, W, H, args.seed, it_per_frame=ITERS)

    fps_plain = 1e3 / ms_plain
    fps_den = 1e3 / ms_den if ms_den else None
    # a denoised frame under 40 dB falls back to the plain full-budget render:
    # charged one plain frame more (frames in flight as timed)
    fallbacks = judge["denoised"]["below_40db"] if judge and judge["denoised"] else None
    eff_ms_den = None if ms_den is None or fallbacks is None else ms_den + fallbacks / args.steps * ms_plain
    den_ok = eff_ms_den is not None
    value = 1e3 / eff_ms_den if den_ok else fps_plain
    line = {
        "metric": "1080p frames/sec at PSNR>=40dB vs converged plain render; iterations/sec",
        "value": value,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": eff_ms_den if den_ok else ms_plain,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded flame (SplitMix64 seed %d)" % args.seed,
        "config": {"workload": (f"denoised render 1920x1080: {pow2(args.low_iters)} iterations, log-density "
                                "tonemap of the histogram, U-Net; "
                                "judged per frame against the 2^32-iteration plain render (PSNR >= 40 dB, else the "
                                "frame falls back to the plain render: 2^30 iterations, DE radius 9)"
                                if den_ok else
                                "plain render 1920x1080, 2^30 iterations, DE radius 9, log-density tonemap") +
                               "; flame: 8 xforms x 2-4 variations + final xform, vibrancy 1",
                   "width": W, "height": H, "iterations": ITERS, "walkers": args.walkers,
                   "threads_per_block": args.tpb, "atomic_width": args.atomic_width, "bin_tile_log2": args.tile,
                   "hist_bits": args.hist_bits, "binning": args.binning,
                   "path": "denoised" if den_ok else "plain", "frames_in_flight": args.pipeline},
        "paths": {"plain": {"fps": fps_plain, "ms": ms_plain, "sequential_ms": seq_plain, "iterations": ITERS,
                            "psnr_db": psnr_plain},
                  "denoised": None if fps_den is None else
                  {"fps": fps_den, "ms": ms_den, "sequential_ms": seq_den, "iterations": args.low_iters,
                   "walkers": args.low_walkers, "fuse": args.low_fuse, "de_radius": args.low_de_radius,
                   "threads_per_block": args.den_tpb, "binning": args.den_binning, "bin_tile_log2": den_tile,
                   "hist_bits": 32,
                   "psnr_db": psnr_den, "fallback_frames": fallbacks,
                   "fps_with_fallback": None if eff_ms_den is None else 1e3 / eff_ms_den,
                   "speedup_vs_plain": (1e3 / eff_ms_den if eff_ms_den else fps_den) / fps_plain,
                   "weights": str(weights.name)}},
        "judge": judge,
        "held_out": held_out,
        "iterations_per_s": it_per_s,
        "iterate_kernel_ms": kms,
        "psnr_db_vs_converged": psnr_den if den_ok else psnr_plain,
        "roofline_denoiser": unet,
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic_bytes, "traffic_measurement": traffic,
                     "kernel": "rfk_iterate" + (" + queue sort/bin kernels" if queue else ""),
                     "unit_of_work": "one bin call of 2^30 iterations", "bytes_per_draw": 16,
                     "draws_per_launch": dr, "launch_ms": kms,
                     "atomic_peak_increments_per_s": peak_inc},
        "cpu_baseline": cpu,
        "autotune": tuned,
    }
    print(json.dumps(line), flush=True)
    return 0


def pow2(n: int) -> str:
    return f"2^{n.bit_length() - 1}" if n > 0 and n & (n - 1) == 0 else str(n)


def score(img, ref) -> dict:
    """PSNR (MAX 1, MSE over RGB of every pixel) and relative L1 (sum |a - b|
    / sum |b| over RGB) of a tonemapped frame against the converged render"""
    import numpy as np
    a = img[..., :3].astype(np.float64)
    b = ref[..., :3].astype(np.float64)
    return {"psnr_db": psnr(a, b), "rel_l1": float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30))}


def summarize(frames: list) -> dict:
    p = [f["psnr_db"] for f in frames]
    r = [f["rel_l1"] for f in frames]
    return {"frames": len(frames), "psnr_db": p, "rel_l1": r, "psnr_db_min": min(p), "psnr_db_mean": sum(p) / len(p),
            "rel_l1_max": max(r), "rel_l1_mean": sum(r) / len(r), "below_40db": sum(1 for v in p if v < 40.0)}


HELD_OUT_SEEDS = range(10001, 10017)  # the denoiser's test flames


def held_out_flames(args, db, lanes, plain, denoised, score, summarize) -> dict:
    """the denoised path on flames the network never saw (1080p flames of
    seeds 10001-10016, the bench's launch shape), each judged against its own
    2^32-iteration plain render"""
    import numpy as np
    from refraktor_amd import _lib
    from refraktor_amd import flame as fl
    W, H = 1920, 1080
    lib = _lib.lib()
    queue = args.binning == "queue"
    frames = []
    for s in HELD_OUT_SEEDS:
        f = fl.flame_1080(s)
        kern = db.compile(f, threads_per_block=args.den_tpb, atomic_width=64, hist_bits=32,
                          bin_queue=args.den_binning == "queue",
                          bin_tile_log2=0 if args.den_binning == "queue" else args.den_tile)
        k64 = db.compile(f, threads_per_block=args.tpb, bin_tile_log2=args.tile, hist_bits=64, bin_queue=queue)
        smp = kern.pack_samples(W, H)
        tone = (f["gamma"], f["brightness"], f["vibrancy"])
        _, _, rb = plain(lanes[0], s + 5000, k64, 1 << 32, _lib.PIXEL_FLOAT4, smp=smp, tone=tone)
        ref = rb.download(np.float32, (H, W, 4))
        _, _, img = denoised(lanes[0], s, kern, smp, tone)
        lib.rfk_synchronize()
        sc = score(img.download(np.float32, (H, W, 3)), ref)
        _, _, lo = plain(lanes[0], s, kern, args.low_iters, _lib.PIXEL_FLOAT4, args.low_walkers, 9.0, smp, tone,
                         args.low_fuse)
        lib.rfk_synchronize()
        sc["plain_same_budget_psnr_db"] = score(lo.download(np.float32, (H, W, 4)), ref)["psnr_db"]
        sc["seed"] = s
        frames.append(sc)
        # the states of these flames are not needed again
        for key in [key for key in lanes[0].states if key[0] in (id(kern), id(k64))]:
            del lanes[0].states[key]
    out = summarize(frames)
    out["seeds"] = [f["seed"] for f in frames]
    out["plain_same_budget_psnr_db"] = [f["plain_same_budget_psnr_db"] for f in frames]
    return out


def shape_args(args) -> list[str]:
    return ["--binning", args.binning, "--tpb", str(args.tpb), "--tile", str(args.tile), "--atomic-width",
            str(args.atomic_width), "--walkers", str(args.walkers), "--hist-bits", str(args.hist_bits),
            "--seed", str(args.seed)]


def run_child(extra: list[str], timeout: float = 600.0) -> dict:
    """python3 bench.py <extra> as a child process; its last stdout line as JSON"""
    import subprocess
    r = subprocess.run([sys.executable, "-u", str(ROOT / "bench.py"), *extra], capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"bench child {extra} failed ({r.returncode}):\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def traffic_probe(args, db, f, W, H, iters) -> int:
    """under rocprofv3 --pmc: warmup, a 1 GiB copy (marker and calibration),
    one bin call of the timed shape, two more copies"""
    from refraktor_amd import _lib, render
    lib = _lib.lib()
    queue = args.binning == "queue"
    k = db.compile(f, threads_per_block=args.tpb, atomic_width=args.atomic_width,
                   bin_tile_log2=0 if queue else args.tile, hist_bits=args.hist_bits, bin_queue=queue)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=args.seed, fuse=20, walkers=args.walkers)
    n = 1 << 30
    src, dst = render.DeviceBuffer(n), render.DeviceBuffer(n)
    src.zero()
    _lib.check(lib.rfk_synchronize())
    _lib.check(lib.rfk_memcpy_dtod(dst.ptr, src.ptr, n))
    _lib.check(lib.rfk_synchronize())
    res = st.bin(iters=iters)
    for _ in range(2):
        _lib.check(lib.rfk_memcpy_dtod(dst.ptr, src.ptr, n))
    _lib.check(lib.rfk_synchronize())
    print(json.dumps({"draws": res["total_draws"], "copy_bytes": n}), flush=True)
    return 0


def _counter_rows(path: Path) -> list[tuple[int, str, float]]:
    """(dispatch, kernel, value) rows of a rocprofv3 counter-collection csv"""
    import csv
    rows = []
    for fpath in path.rglob("*counter_collection.csv"):
        with open(fpath, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def live_traffic(args) -> dict:
    """HBM bytes of one bin call of the timed launch shape: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) over the dispatches between the probe's
    first and second 1 GiB copy, scaled by what the same counters read for
    those copies (bytes per KiB)"""
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return {"note": "rocprofv3 not found: traffic not measured"}
    out = {"shape": {"binning": args.binning, "threads_per_block": args.tpb, "walkers": args.walkers,
                     "hist_bits": args.hist_bits, "atomic_width": args.atomic_width}}
    with tempfile.TemporaryDirectory() as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = Path(tmp) / counter
            cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", str(d), "-o", "run", "--",
                   sys.executable, str(ROOT / "bench.py"), "--traffic-probe", *shape_args(args)]
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
            except subprocess.TimeoutExpired:
                return {**out, "note": f"rocprofv3 {counter} run timed out"}
            lines = r.stdout.strip().splitlines()
            probe = None
            for line in reversed(lines):
                if line.startswith("{"):
                    probe = json.loads(line)
                    break
            rows = _counter_rows(d)
            # rocprofv3 can exit with an error after a complete run: the
            # probe's own last line and the counter file decide
            if probe is None or not rows:
                return {**out, "note": f"rocprofv3 {counter} run gave no counters (exit {r.returncode})"}
            # the probe's 1 GiB copies (the bin call makes small copies of its own)
            copies = [i for i, (_, name, v) in enumerate(rows) if "copyBuffer" in name and v > 65536.0]
            if len(copies) != 3:
                return {**out, "note": f"{counter}: the probe's copies were not found"}
            kib_copy = [rows[i][2] for i in copies[:3]]
            scale = probe["copy_bytes"] / (sum(kib_copy) / 3.0)
            kib_bin = sum(v for _, name, v in rows[copies[0] + 1:copies[1]])
            key = "read" if counter == "FETCH_SIZE" else "write"
            out[f"{key}_bytes"] = kib_bin * scale
            out[f"{key}_scale_bytes_per_kib"] = scale
            out["draws"] = probe["draws"]
    out["bytes_per_bin_call"] = out["read_bytes"] + out["write_bytes"]
    out["bytes_per_draw"] = out["bytes_per_bin_call"] / out["draws"]
    out["method"] = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over the dispatches of one bin call of this "
                     "shape in a child process, each scaled by the counter's reading of a 1 GiB device copy")
    return out


def cpu_baseline(f: dict, W: int, H: int, seed: int, it_per_frame: int) -> dict:
    """the oracle's OpenMP render of the same flame in single precision,
    compiled here for this host (-O3 -march=native), on a bounded sample"""
    import tempfile
    from oracle import build as obuild
    from oracle import oracle
    from refraktor_amd import render
    blocks, tpb = 1024, 256
    passes = 1024  # 1024*256*1024 = 2^28 iterations
    L, P32, pals = oracle.prepare(f, render.CONFIG_DIR, W, H, blocks, seed)
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)
    model = "unknown"
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    with tempfile.TemporaryDirectory() as tmp:
        lib = obuild.build_f32(Path(tmp) / "liboracle_f32.so")
        oracle.render_f32(lib, L, P32, P32.shape[1], pals, 64, tpb, W, H, seed, 20, 16, threads=threads)  # warm
        t0 = time.perf_counter()
        _, draws = oracle.render_f32(lib, L, P32, P32.shape[1], pals, blocks, tpb, W, H, seed, 20, passes,
                                     threads=threads)
        dt = time.perf_counter() - t0
    iters = blocks * tpb * passes
    ips = iters / dt
    return {"value": ips / it_per_frame, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{iters} iterations (+20 fuse passes) of the same 1080p flame, single precision, "
                      f"gcc -O3 -march=native -fopenmp, {threads} threads on {model} ({os.cpu_count()} "
                      f"logical CPUs visible) in {dt:.2f} s ({ips:.3e} it/s), scaled to a 2^30-iteration frame",
            "iterations_per_s": ips, "cpu_model": model}


if __name__ == "__main__":
    sys.exit(main())

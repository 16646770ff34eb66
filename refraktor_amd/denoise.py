"""Learned denoiser: data generation with the renderer, GPU training, inference.

Training pairs come from the renderer itself: for a seeded random flame the
low-budget frame (features of its histogram and tonemapped image) is paired
with the tonemapped converged frame of the same flame. Both go through the
plain pipeline (density estimation + tonemap). Training and inference run the
library's own kernels (rfk_denoiser_*); this module only orchestrates.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from . import flame as fl
from . import render
from ._lib import check

WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "unet_1080.bin"
FEAT_C = 8


class Denoiser:
    """rfkt::denoiser counterpart (interface/denoiser.h:21-39) backed by the U-Net."""

    def __init__(self, seed: int = 1, weights: str | Path | None = None):
        self._lib = _lib.lib()
        h = C.c_void_p()
        check(self._lib.rfk_denoiser_create(seed, C.byref(h)))
        self.handle = h.value
        if weights is not None:
            self.load(weights)

    def param_count(self) -> tuple[int, int]:
        nw, nb = C.c_size_t(), C.c_size_t()
        check(self._lib.rfk_denoiser_param_count(self.handle, C.byref(nw), C.byref(nb)))
        return nw.value, nb.value

    def load(self, path) -> None:
        check(self._lib.rfk_denoiser_load(self.handle, str(path).encode()))

    def save(self, path) -> None:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        check(self._lib.rfk_denoiser_save(self.handle, str(path).encode()))

    def params(self) -> tuple[np.ndarray, np.ndarray]:
        nw, nb = self.param_count()
        w, b = np.empty(nw, np.float32), np.empty(nb, np.float32)
        check(self._lib.rfk_denoiser_get_params(self.handle, w.ctypes.data, b.ctypes.data))
        return w, b

    def set_params(self, w: np.ndarray, b: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        check(self._lib.rfk_denoiser_set_params(self.handle, w.ctypes.data, b.ctypes.data))

    def grads(self) -> tuple[np.ndarray, np.ndarray]:
        nw, nb = self.param_count()
        w, b = np.empty(nw, np.float32), np.empty(nb, np.float32)
        check(self._lib.rfk_denoiser_get_grads(self.handle, w.ctypes.data, b.ctypes.data))
        return w, b

    def forward(self, feat: render.DeviceBuffer, base: render.DeviceBuffer, n: int, h: int, w: int,
                out: render.DeviceBuffer, stream=None) -> None:
        """enqueued on `stream` (a state's stream keeps the frame's work in order)"""
        check(self._lib.rfk_denoiser_forward(self.handle, feat.ptr, base.ptr, n, h, w, out.ptr, stream))

    def denoise(self, image: render.DeviceBuffer, in_format: int, out: render.DeviceBuffer, out_format: int,
                width: int, height: int, stream=None) -> None:
        """denoiser::denoise (interface/denoiser.h:35-38): a tonemapped device
        image (float4 / half4 / float3 / half3) in, the denoised one out,
        enqueued on `stream`"""
        check(self._lib.rfk_denoise(self.handle, image.ptr, in_format, out.ptr, out_format, width, height, stream))

    def train_step(self, feat, base, target, n: int, h: int, w: int, lr: float) -> float:
        loss = C.c_float()
        check(self._lib.rfk_denoiser_train_step(self.handle, feat.ptr, base.ptr, target.ptr, n, h, w, lr,
                                                C.byref(loss), None))
        return loss.value

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_denoiser_free(self.handle)
        except Exception:
            pass


@dataclass
class Frame:
    """one rendered frame kept on the device"""
    hist: render.DeviceBuffer     # RGBA fixed-point bins of hist_bits
    hist_bits: int
    tm: render.DeviceBuffer       # float4 tonemapped (after DE)
    width: int
    height: int
    quality: float


def render_frame(db: render.VariationDB, f: dict, width: int, height: int, iters: int, seed: int,
                 de_radius: float, walkers: int = 1 << 20, kernel=None) -> Frame:
    """plain pipeline on the device: render, density estimation, tonemap (float4)"""
    lib = _lib.lib()
    k = kernel or db.compile(f)
    st = k.warmup(k.pack_samples(width, height), width, height, seed=seed, fuse=20,
                  walkers=min(walkers, max(iters // 16, 1024)))
    st.bin(iters=iters)
    inf = st.info()
    npx = width * height
    hb = npx * 4 * inf.hist_bits // 8
    hist = render.DeviceBuffer(hb)
    check(lib.rfk_memcpy_dtod(hist.ptr, inf.histogram, hb))
    de = render.DeviceBuffer(npx * 16)
    check(lib.rfk_density_estimate(st.handle, de_radius, 0.0, 0.4, de.ptr))
    tm = render.DeviceBuffer(npx * 16)
    args = _lib.TonemapArgs(inf.quality, f["gamma"], f["brightness"], f["vibrancy"], 0.0)
    # density estimation runs on the state's stream: tonemap behind it there,
    # and hand over the frame only once both are done
    check(lib.rfk_tonemap_image(de.ptr, width, height, tm.ptr, _lib.PIXEL_FLOAT4, C.byref(args), inf.stream))
    check(lib.rfk_synchronize())
    return Frame(hist, inf.hist_bits, tm, width, height, inf.quality)


def features(frame: Frame) -> tuple[render.DeviceBuffer, render.DeviceBuffer]:
    npx = frame.width * frame.height
    feat = render.DeviceBuffer(npx * FEAT_C * 2)
    base = render.DeviceBuffer(npx * 12)
    check(_lib.lib().rfk_denoiser_features(frame.hist.ptr, frame.hist_bits, frame.tm.ptr, npx, feat.ptr, base.ptr,
                                           None))
    return feat, base


def rgb_of(frame: Frame) -> render.DeviceBuffer:
    """the frame's tonemapped RGB as FP32 (N*H*W*3), via the feature kernel's base output"""
    return features(frame)[1]


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(1.0 / mse)


@dataclass
class Dataset:
    feat: render.DeviceBuffer    # [n][H][W][8] bf16
    base: render.DeviceBuffer    # [n][H][W][3] f32
    target: render.DeviceBuffer  # [n][H][W][3] f32
    n: int
    height: int
    width: int


def make_pairs(seeds, width: int, height: int, low_iters: int, high_iters: int, de_radius: float,
               log=print) -> Dataset:
    """render (low, converged) pairs of random flames; all data stays in HBM"""
    db = render.VariationDB()
    n = len(seeds)
    npx = width * height
    feat = render.DeviceBuffer(n * npx * FEAT_C * 2)
    base = render.DeviceBuffer(n * npx * 12)
    tgt = render.DeviceBuffer(n * npx * 12)
    lib = _lib.lib()
    t0 = time.time()
    for i, s in enumerate(seeds):
        f = fl.flame_keyframe(s)
        k = db.compile(f)
        lo = render_frame(db, f, width, height, low_iters, seed=s * 7 + 1, de_radius=de_radius, kernel=k)
        hi = render_frame(db, f, width, height, high_iters, seed=s * 7 + 2, de_radius=de_radius, kernel=k)
        fe, ba = features(lo)
        t = rgb_of(hi)
        check(lib.rfk_memcpy_dtod(feat.ptr + i * npx * FEAT_C * 2, fe.ptr, npx * FEAT_C * 2))
        check(lib.rfk_memcpy_dtod(base.ptr + i * npx * 12, ba.ptr, npx * 12))
        check(lib.rfk_memcpy_dtod(tgt.ptr + i * npx * 12, t.ptr, npx * 12))
        if log and (i % 16 == 0 or i == n - 1):
            log(f"[pairs] {i + 1}/{n} flames rendered, {time.time() - t0:.1f} s")
    return Dataset(feat, base, tgt, n, height, width)


def _crop_offsets(rng, data: Dataset, batch: int, crop: int, importance) -> np.ndarray:
    """(image, y, x) of `batch` crops; with `importance` ([n][gy][gx] coarse
    weights over 16-pixel cells) crops favour bright, detailed regions"""
    if importance is None:
        return np.stack([rng.integers(0, data.n, batch), rng.integers(0, data.height - crop + 1, batch),
                         rng.integers(0, data.width - crop + 1, batch)], axis=1).astype(np.int32)
    out = np.empty((batch, 3), np.int32)
    for b in range(batch):
        i = int(rng.integers(0, data.n))
        w = importance[i].ravel()
        cell = int(rng.choice(w.size, p=w / w.sum()))
        cy, cx = divmod(cell, importance[i].shape[1])
        y = int(np.clip(cy * 16 + 8 - crop // 2 + rng.integers(-crop // 4, crop // 4 + 1), 0, data.height - crop))
        x = int(np.clip(cx * 16 + 8 - crop // 2 + rng.integers(-crop // 4, crop // 4 + 1), 0, data.width - crop))
        out[b] = (i, y, x)
    return out


def train(den: Denoiser, data: Dataset, steps: int, batch: int = 8, crop: int = 128, lr: float = 1e-3,
          seed: int = 0, log=print, importance=None) -> list[float]:
    """Adam on random crops of the dataset (cosine learning-rate decay)"""
    rng = np.random.default_rng(seed)
    lib = _lib.lib()
    cp = crop * crop
    bf = render.DeviceBuffer(batch * cp * FEAT_C * 2)
    bb = render.DeviceBuffer(batch * cp * 12)
    bt = render.DeviceBuffer(batch * cp * 12)
    losses = []
    t0 = time.time()
    for step in range(steps):
        offs = _crop_offsets(rng, data, batch, crop, importance)
        op = offs.ctypes.data
        check(lib.rfk_nn_gather_crops(data.feat.ptr, data.n, data.height, data.width, FEAT_C * 2, op, batch, crop, crop,
                                      bf.ptr, None))
        check(lib.rfk_nn_gather_crops(data.base.ptr, data.n, data.height, data.width, 12, op, batch, crop, crop, bb.ptr,
                                      None))
        check(lib.rfk_nn_gather_crops(data.target.ptr, data.n, data.height, data.width, 12, op, batch, crop, crop, bt.ptr,
                                      None))
        cur_lr = lr * (0.5 * (1 + np.cos(np.pi * step / max(steps, 1))))
        losses.append(den.train_step(bf, bb, bt, batch, crop, crop, float(cur_lr)))
        if log and (step % 100 == 0 or step == steps - 1):
            log(f"[train] step {step} loss {np.mean(losses[-100:]):.3e} ({time.time() - t0:.1f} s)")
    return losses


def denoise_frame(den: Denoiser, frame: Frame) -> np.ndarray:
    """denoised RGB (host float32, H x W x 3)"""
    feat, base = features(frame)
    out = render.DeviceBuffer(frame.width * frame.height * 12)
    den.forward(feat, base, 1, frame.height, frame.width, out)
    return out.download(np.float32, (frame.height, frame.width, 3))

"""GPU JPEG encoder (rfkt::jpeg_encoder: interface/jpeg_encoder.h:11-24)
against the oracle's restatement: the same JFIF bytes, file for file. The
oracle itself is pinned by PIL's decoder in test_cpu_oracle.py."""
import io

import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import _lib
from refraktor_amd import flame as fl
from refraktor_amd import render

pytestmark = pytest.mark.gpu


def _image(w: int, h: int, seed: int, channels: int = 3) -> np.ndarray:
    """a gradient with a noise patch, flat areas and saturated primaries, so
    blocks run from all-zero to full 10-bit AC coefficients"""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([x * 255 // max(w, 1), y * 255 // max(h, 1), (x ^ y) & 255], -1).astype(np.uint8)
    img[: h // 2, : w // 3] = rng.integers(0, 256, size=(h // 2, w // 3, 3), dtype=np.uint8)
    img[h // 2:, w // 3: w // 2] = (0, 0, 255)
    img[h // 2:, w // 2: 2 * w // 3] = (255, 0, 0)
    img[: h // 3, 2 * w // 3:] = 128
    if channels == 4:
        img = np.concatenate([img, rng.integers(0, 256, size=(h, w, 1), dtype=np.uint8)], -1)
    return img


@pytest.mark.parametrize("w,h", [(1, 1), (13, 11), (70, 45), (83, 61), (640, 360), (1920, 1080)])
@pytest.mark.parametrize("quality", [1, 50, 90, 100])
def test_jpeg_matches_oracle(gpu, w, h, quality):
    img = _image(w, h, w * 7 + h + quality)
    enc = render.JpegEncoder()
    assert enc.encode(img, quality) == oracle.jpeg_encode(img, quality)


def test_jpeg_rgba_pitch_and_reuse(gpu):
    """RGBA input read through a padded row pitch; one encoder reused across
    sizes (its buffers grow and are reused) gives each file afresh"""
    enc = render.JpegEncoder()
    big, small = _image(333, 217, 5, 4), _image(40, 24, 6, 4)
    for img, q in [(big, 85), (small, 20), (big, 60)]:
        h, w, _ = img.shape
        pitch = w * 4 + 36
        padded = np.zeros((h, pitch), np.uint8)
        padded[:, : w * 4] = img.reshape(h, w * 4)
        buf = render.DeviceBuffer(padded.nbytes)
        buf.upload(padded)
        got = enc.encode_device(buf.ptr, w, h, q, _lib.PIXEL_RGBA8, pitch=pitch)
        assert got == oracle.jpeg_encode(img, q)


def test_jpeg_rejects_bad_arguments(gpu):
    lib = _lib.lib()
    enc = render.JpegEncoder()
    buf = render.DeviceBuffer(64 * 64 * 4)
    n = _lib.C.c_size_t()
    assert lib.rfk_jpeg_retrieve(enc.handle, None, 0, _lib.C.byref(n)) != 0  # nothing encoded yet
    assert lib.rfk_jpeg_encode(enc.handle, buf.ptr, _lib.PIXEL_HALF4, 64 * 8, 64, 64, 90, None) != 0
    assert lib.rfk_jpeg_encode(enc.handle, buf.ptr, _lib.PIXEL_RGB8, 64 * 3, 0, 64, 90, None) != 0
    assert lib.rfk_jpeg_encode(enc.handle, buf.ptr, _lib.PIXEL_RGB8, 64 * 3, 64, 64, 0, None) != 0
    assert lib.rfk_jpeg_encode(enc.handle, buf.ptr, _lib.PIXEL_RGB8, 10, 64, 64, 90, None) != 0  # pitch < row
    # quality above 100 encodes as 100
    img = _image(64, 64, 1)
    assert enc.encode(img, 250) == oracle.jpeg_encode(img, 100)


def test_render_to_jpeg(gpu, db):
    """the output path of a frame: histogram -> tonemap (half4) -> 8-bit RGB
    conversion -> JPEG, the file equal to the oracle's and decoded by PIL
    close to the 8-bit frame"""
    from PIL import Image
    W, H = 320, 180
    k = db.compile(fl.flame_small(2))
    st = k.warmup(k.pack_samples(W, H), W, H, seed=3)
    st.bin(iters=1 << 22)
    half = st.tonemap(2.2, 4.0, 1.0, fmt=_lib.PIXEL_HALF4)
    rgb = render.convert_image(half, _lib.PIXEL_RGB8)
    data = render.JpegEncoder().encode(rgb, 90)
    assert data == oracle.jpeg_encode(rgb, 90)
    dec = np.asarray(Image.open(io.BytesIO(data)).convert("RGB")).astype(np.float64)
    mse = ((dec - rgb) ** 2).mean()
    assert 10 * np.log10(255.0 ** 2 / max(mse, 1e-9)) > 30.0

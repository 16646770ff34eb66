"""rfk_denoise, the reference's denoise call (interface/denoiser.h:35-38): a
tonemapped image in, the denoised image out, in the four pixel formats, at
any size; against the histogram-feature path the network was trained on."""
import ctypes as C

import numpy as np
import pytest

from refraktor_amd import _lib
from refraktor_amd import denoise
from refraktor_amd import flame as fl
from refraktor_amd import render

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not denoise.DEFAULT_WEIGHTS.exists(), reason="no trained weights in the package")]

W, H = 1920, 1080


@pytest.fixture(scope="module")
def frame(db, gpu):
    f = fl.flame_1080(1)
    k = db.compile(f, threads_per_block=1024, atomic_width=64, bin_tile_log2=12, hist_bits=64)
    lo = denoise.render_frame(db, f, W, H, 1 << 25, seed=3, de_radius=0.0, kernel=k)
    hi = denoise.render_frame(db, f, W, H, 1 << 32, seed=4, de_radius=9.0, kernel=k)
    den = denoise.Denoiser(weights=denoise.DEFAULT_WEIGHTS)
    hist_path = denoise.denoise_frame(den, lo)
    ref = denoise.rgb_of(hi).download(np.float32, (H, W, 3))
    tm = lo.tm.download(np.float32, (H, W, 4))
    return den, lo, tm, hist_path, ref


def _upload(a: np.ndarray) -> render.DeviceBuffer:
    b = render.DeviceBuffer(a.nbytes)
    b.upload(np.ascontiguousarray(a))
    return b


def _run(den, img: np.ndarray, fmt_in: int, fmt_out: int, w: int, h: int) -> np.ndarray:
    ch_out = 4 if fmt_out in (_lib.PIXEL_FLOAT4, _lib.PIXEL_HALF4) else 3
    dt_out = np.float32 if fmt_out in (_lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT3) else np.float16
    src = _upload(img)
    out = render.DeviceBuffer(w * h * ch_out * np.dtype(dt_out).itemsize)
    den.denoise(src, fmt_in, out, fmt_out, w, h)
    _lib.check(_lib.lib().rfk_synchronize())
    return out.download(dt_out, (h, w, ch_out)).astype(np.float32)


def test_float4_image_matches_histogram_path(frame):
    """the alpha channel carries the density: from the float4 tonemap alone
    the call reproduces the histogram-feature path"""
    den, lo, tm, hist_path, ref = frame
    out = _run(den, tm, _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT4, W, H)
    assert np.array_equal(out[..., 3], tm[..., 3])  # alpha passes through
    p = denoise.psnr(out[..., :3], hist_path)
    print(f"[denoise api] float4 vs histogram path {p:.1f} dB; vs converged {denoise.psnr(out[..., :3], ref):.2f} dB")
    assert p > 60.0


def test_half_and_three_channel_formats(frame):
    den, lo, tm, hist_path, ref = frame
    full = _run(den, tm, _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT4, W, H)[..., :3]
    # half4 in: the half-precision alpha still gives the density
    h4 = _run(den, tm.astype(np.float16), _lib.PIXEL_HALF4, _lib.PIXEL_HALF4, W, H)
    assert denoise.psnr(h4[..., :3], full) > 45.0
    # output formats of one input agree (half rounding aside)
    f3 = _run(den, tm, _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT3, W, H)
    np.testing.assert_array_equal(f3, full)
    hh3 = _run(den, tm, _lib.PIXEL_FLOAT4, _lib.PIXEL_HALF3, W, H)
    np.testing.assert_allclose(hh3, full, atol=1e-3)
    # 3-channel input: no density features, still a gain over the raw frame
    rgb_only = _run(den, np.ascontiguousarray(tm[..., :3]), _lib.PIXEL_FLOAT3, _lib.PIXEL_FLOAT3, W, H)
    h3 = _run(den, tm[..., :3].astype(np.float16), _lib.PIXEL_HALF3, _lib.PIXEL_FLOAT3, W, H)
    assert denoise.psnr(h3, rgb_only) > 45.0
    p_raw, p3, p4 = denoise.psnr(tm[..., :3], ref), denoise.psnr(rgb_only, ref), denoise.psnr(full, ref)
    print(f"[denoise api] vs converged: raw {p_raw:.2f} dB, 3-channel {p3:.2f} dB, 4-channel {p4:.2f} dB")
    assert np.isfinite(rgb_only).all() and rgb_only.min() >= 0.0 and rgb_only.max() <= 1.0
    assert p3 > p_raw


def test_any_size(frame):
    """sizes that are not multiples of 8 are padded and cropped back: away
    from the new border (beyond the network's reach) the result is the
    full-size one"""
    den, lo, tm, hist_path, ref = frame
    full = _run(den, tm, _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT4, W, H)
    w, h = 1917, 1075
    crop = np.ascontiguousarray(tm[:h, :w])
    out = _run(den, crop, _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT4, w, h)
    m = 64
    np.testing.assert_allclose(out[m:h - m, m:w - m], full[m:h - m, m:w - m], atol=1e-6)
    small = _run(den, np.ascontiguousarray(tm[500:533, 900:931]), _lib.PIXEL_FLOAT4, _lib.PIXEL_FLOAT4, 31, 33)
    assert np.isfinite(small).all()


def test_bad_arguments_and_allocation_failure(frame):
    den, lo, tm, hist_path, ref = frame
    lib = _lib.lib()
    src = _upload(tm)
    out = render.DeviceBuffer(W * H * 16)
    assert lib.rfk_denoise(den.handle, src.ptr, _lib.PIXEL_RGBA8, out.ptr, _lib.PIXEL_FLOAT4, W, H, None) \
        == _lib.RFK_ERR_INVALID
    assert lib.rfk_denoise(den.handle, src.ptr, _lib.PIXEL_FLOAT4, src.ptr, _lib.PIXEL_FLOAT4, W, H, None) \
        == _lib.RFK_ERR_INVALID
    assert lib.rfk_denoise(den.handle, src.ptr, _lib.PIXEL_FLOAT4, out.ptr, _lib.PIXEL_FLOAT4, 0, H, None) \
        == _lib.RFK_ERR_INVALID
    # a failed activation allocation is reported (nothing is launched) and
    # the next call, at a shape not seen before, allocates and runs
    w, h = 1000, 600
    crop = _upload(np.ascontiguousarray(tm[:h, :w]))
    o2 = render.DeviceBuffer(w * h * 16)
    assert lib.rfk_nn_inject_alloc_failures(1) == 0
    assert lib.rfk_denoise(den.handle, crop.ptr, _lib.PIXEL_FLOAT4, o2.ptr, _lib.PIXEL_FLOAT4, w, h, None) \
        == _lib.RFK_ERR_HIP
    assert b"allocation" in lib.rfk_last_error()
    assert lib.rfk_nn_inject_alloc_failures(3) == 0
    assert lib.rfk_denoise(den.handle, crop.ptr, _lib.PIXEL_FLOAT4, o2.ptr, _lib.PIXEL_FLOAT4, w, h, None) \
        == _lib.RFK_ERR_HIP
    assert lib.rfk_nn_inject_alloc_failures(0) == 0
    _lib.check(lib.rfk_denoise(den.handle, crop.ptr, _lib.PIXEL_FLOAT4, o2.ptr, _lib.PIXEL_FLOAT4, w, h, None))
    _lib.check(lib.rfk_synchronize())
    got = o2.download(np.float32, (h, w, 4))
    np.testing.assert_allclose(got, _run(den, np.ascontiguousarray(tm[:h, :w]), _lib.PIXEL_FLOAT4,
                                         _lib.PIXEL_FLOAT4, w, h), atol=1e-6)

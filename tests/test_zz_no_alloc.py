"""rfk_nn_gather_crops_noalloc (offsets staged in a static device array, no
allocation) against rfk_nn_gather_crops (per-call allocation): the same bytes
on two calls in a row, the second with fewer and smaller crops, and the same
refusals."""
import numpy as np
import pytest

from refraktor_amd import _lib
from refraktor_amd import render

pytestmark = pytest.mark.gpu

N_IMG, H, W, PB = 3, 64, 80, 16


def _offsets(rng, n, ch, cw):
    return np.stack([rng.integers(0, N_IMG, n), rng.integers(0, H - ch + 1, n), rng.integers(0, W - cw + 1, n)],
                    axis=1).astype(np.int32)


def test_noalloc_gather_matches_allocating_call(gpu):
    lib = _lib.lib()
    rng = np.random.default_rng(5)
    stack = rng.integers(0, 256, N_IMG * H * W * PB, dtype=np.uint8)
    src = render.DeviceBuffer(stack.nbytes)
    src.upload(stack)
    for n, ch, cw in ((300, 16, 24), (37, 8, 8)):
        offs = _offsets(rng, n, ch, cw)
        a = render.DeviceBuffer(n * ch * cw * PB)
        b = render.DeviceBuffer(n * ch * cw * PB)
        a.zero()
        b.zero()
        _lib.check(lib.rfk_nn_gather_crops(src.ptr, N_IMG, H, W, PB, offs.ctypes.data, n, ch, cw, a.ptr, None))
        _lib.check(lib.rfk_nn_gather_crops_noalloc(src.ptr, N_IMG, H, W, PB, offs.ctypes.data, n, ch, cw, b.ptr,
                                                   None))
        ga = a.download(np.uint8, (n, ch, cw, PB))
        gb = b.download(np.uint8, (n, ch, cw, PB))
        assert np.array_equal(ga, gb), (n, ch, cw)
        # and both are the crops themselves
        img = stack.reshape(N_IMG, H, W, PB)
        for k in (0, n // 2, n - 1):
            i, y, x = offs[k]
            assert np.array_equal(gb[k], img[i, y:y + ch, x:x + cw])


def test_noalloc_gather_refusals(gpu):
    lib = _lib.lib()
    src = render.DeviceBuffer(N_IMG * H * W * PB)
    dst = render.DeviceBuffer(8 * 8 * PB * 8)
    too_many = np.zeros((4097, 3), np.int32)
    assert lib.rfk_nn_gather_crops_noalloc(src.ptr, N_IMG, H, W, PB, too_many.ctypes.data, 4097, 8, 8, dst.ptr,
                                           None) == _lib.RFK_ERR_INVALID
    bad = np.array([[N_IMG, 0, 0]], np.int32)
    for fn in (lib.rfk_nn_gather_crops, lib.rfk_nn_gather_crops_noalloc):
        assert fn(src.ptr, N_IMG, H, W, PB, bad.ctypes.data, 1, 8, 8, dst.ptr, None) == _lib.RFK_ERR_INVALID

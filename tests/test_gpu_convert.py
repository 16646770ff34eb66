"""rfk_convert_image (the converter after the tonemap) against the oracle,
bit for bit, for every input/output pair and flag combination."""
import itertools

import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import _lib
from refraktor_amd import render

pytestmark = pytest.mark.gpu


def _image(rng, h, w, c):
    v = rng.uniform(-0.1, 1.2, size=(h, w, c))
    v[0, :4, 0] = [0.0, 1.0, 0.5, 1 / 255]  # truncation and clamp edges
    return v.astype(np.float16)


@pytest.mark.parametrize("channels", [3, 4])
def test_convert_matches_oracle(gpu, channels):
    rng = np.random.default_rng(40 + channels)
    img = _image(rng, 37, 61, channels)
    for fmt, flip, swap in itertools.product(range(6), (False, True), (False, True)):
        got = render.convert_image(img, fmt, flip=flip, swap=swap)
        want = oracle.convert_image(img, fmt, flip=flip, swap=swap)
        assert got.shape == want.shape and got.dtype == want.dtype, (fmt, flip, swap)
        if got.dtype == np.float16:
            np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16), err_msg=str((fmt, flip, swap)))
        else:
            np.testing.assert_array_equal(got, want, err_msg=str((fmt, flip, swap)))


def test_convert_rejects_bad_formats(gpu):
    lib = _lib.lib()
    buf = render.DeviceBuffer(64)
    assert lib.rfk_convert_image(buf.ptr, _lib.PIXEL_FLOAT4, buf.ptr, 0, 2, 2, 0, None) != 0
    assert lib.rfk_convert_image(buf.ptr, _lib.PIXEL_HALF4, buf.ptr, 6, 2, 2, 0, None) != 0
    assert lib.rfk_convert_image(buf.ptr, _lib.PIXEL_HALF4, buf.ptr, 0, 2, 2, 4, None) != 0

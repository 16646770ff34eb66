"""Pins the denoiser oracle: finite differences and closed forms."""
import numpy as np

from oracle import nn_oracle as nn


def _fd(f, x, idx, h=1e-6):
    x1, x2 = x.copy(), x.copy()
    x1.flat[idx] += h
    x2.flat[idx] -= h
    return (f(x1) - f(x2)) / (2 * h)


def test_conv_identity_kernel():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 7, 3))
    w = np.zeros((3, 9, 3))
    for c in range(3):
        w[c, 4, c] = 1.0  # centre tap
    np.testing.assert_allclose(nn.conv3x3(x, w), x)
    # a shift kernel moves the image by one pixel with zero fill
    w = np.zeros((3, 9, 3))
    for c in range(3):
        w[c, 5, c] = 1.0  # tap (dy=0, dx=+1)
    y = nn.conv3x3(x, w)
    np.testing.assert_allclose(y[:, :, :-1], x[:, :, 1:])
    assert np.all(y[:, :, -1] == 0)


def test_conv_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 5, 3))
    w = rng.normal(size=(2, 9, 3))
    gy = rng.normal(size=(1, 4, 5, 2))
    loss_w = lambda ww: float(np.sum(nn.conv3x3(x, ww) * gy))  # noqa: E731
    loss_x = lambda xx: float(np.sum(nn.conv3x3(xx, w) * gy))  # noqa: E731
    dw = nn.conv3x3_wgrad(x, gy)
    dx = nn.conv3x3_dgrad(gy, w)
    for idx in (0, 7, 20, 53):
        assert abs(_fd(loss_w, w, idx) - dw.flat[idx]) < 1e-6
    for idx in (0, 11, 33, 59):
        assert abs(_fd(loss_x, x, idx) - dx.flat[idx]) < 1e-6


def test_pool_and_upsample_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(1, 4, 6, 2))
    b = rng.normal(size=(1, 2, 3, 2))
    # <pool(a), b> == <a, pool_bwd(b)> and <up(b), a> == <b, up_bwd(a)>
    assert abs(np.sum(nn.pool(a) * b) - np.sum(a * nn.pool_bwd(b))) < 1e-12
    assert abs(np.sum(nn.up(b) * a) - np.sum(b * nn.up_bwd(a))) < 1e-12


def test_unet_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    shapes = nn.layer_shapes()
    nw = sum(o * 9 * (a + b) for a, b, o, _ in shapes)
    nb = sum(o for _, _, o, _ in shapes)
    w = rng.normal(size=nw) * 0.2
    b = rng.normal(size=nb) * 0.05
    feat = rng.normal(size=(1, 8, 8, nn.FEAT))
    base = rng.uniform(size=(1, 8, 8, 3))
    tgt = rng.uniform(size=(1, 8, 8, 3))
    loss, gw, gb, _ = nn.unet_loss_and_grads(feat, base, tgt, w, b, bf16_acts=False)
    f = lambda ww: nn.unet_loss_and_grads(feat, base, tgt, ww, b, bf16_acts=False)[0]  # noqa: E731
    idxs = rng.choice(nw, 12, replace=False)
    for i in idxs:
        fd = _fd(f, w, int(i), h=1e-5)
        assert abs(fd - gw[i]) <= 1e-6 + 1e-3 * abs(fd), (i, fd, gw[i])
    fb = lambda bb: nn.unet_loss_and_grads(feat, base, tgt, w, bb, bf16_acts=False)[0]  # noqa: E731
    for i in rng.choice(nb, 6, replace=False):
        fd = _fd(fb, b, int(i), h=1e-5)
        assert abs(fd - gb[i]) <= 1e-6 + 1e-3 * abs(fd)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.0078125, 3.14159265, -2.5e-3], dtype=np.float32)
    r = nn.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 is a tie: rounds to even
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] - 3.140625) < 1e-7
    assert np.all(nn.from_bf16_bits(nn.bf16_bits(v)) == r)

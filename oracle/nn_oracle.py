"""ORACLE — test infrastructure only: float64 numpy restatement of the
denoiser's layers and of the whole U-Net with its gradients.

Layout: activations NHWC, weights [Cout][9][Cin] with tap = 3*dy + dx over
offsets (dy-1, dx-1), "same" zero padding. The U-Net matches the layer list
of the library's denoiser: enc 32/64/96/128 with 2x2 average pooling, nearest
upsampling, skip concatenation [upsampled, skip], ReLU everywhere except the
output conv, output = base + residual[..., :3], loss = mean squared error.
"""
from __future__ import annotations

import numpy as np

CH = (32, 64, 96, 128)
FEAT, OUT = 8, 8


def to_bf16(a: np.ndarray) -> np.ndarray:
    """round float32 to bfloat16 (nearest even), returned as float32"""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits(a: np.ndarray) -> np.ndarray:
    return (to_bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(x: np.ndarray, w: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    N, H, W, C = x.shape
    Cout = w.shape[0]
    xp = _pad(x.astype(np.float64))
    y = np.zeros((N, H, W, Cout))
    for tap in range(9):
        dy, dx = divmod(tap, 3)
        y += xp[:, dy:dy + H, dx:dx + W, :] @ w[:, tap, :].T.astype(np.float64)
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(x: np.ndarray, dy: np.ndarray) -> np.ndarray:
    N, H, W, C = x.shape
    Cout = dy.shape[-1]
    xp = _pad(x.astype(np.float64))
    dw = np.zeros((Cout, 9, C))
    g = dy.reshape(-1, Cout).astype(np.float64)
    for tap in range(9):
        ddy, ddx = divmod(tap, 3)
        dw[:, tap, :] = g.T @ xp[:, ddy:ddy + H, ddx:ddx + W, :].reshape(-1, C)
    return dw


def conv3x3_dgrad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    N, H, W, Cout = dy.shape
    Cin = w.shape[2]
    dx = np.zeros((N, H + 2, W + 2, Cin))
    for tap in range(9):
        ddy, ddx = divmod(tap, 3)
        dx[:, ddy:ddy + H, ddx:ddx + W, :] += dy.astype(np.float64) @ w[:, tap, :].astype(np.float64)
    return dx[:, 1:-1, 1:-1, :]


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_bwd(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(g):
    N, H, W, C = g.shape
    return g.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def layer_shapes():
    c0, c1, c2, c3 = CH
    return [(FEAT, 0, c0, 1), (c0, 0, c0, 1), (c0, 0, c1, 1), (c1, 0, c1, 1), (c1, 0, c2, 1), (c2, 0, c2, 1),
            (c2, 0, c3, 1), (c3, 0, c3, 1), (c3, c2, c2, 1), (c2, 0, c2, 1), (c2, c1, c1, 1), (c1, 0, c1, 1),
            (c1, c0, c0, 1), (c0, 0, c0, 1), (c0, 0, OUT, 0)]


def split_params(wflat: np.ndarray, bflat: np.ndarray):
    ws, bs, wo, bo = [], [], 0, 0
    for a, b, o, _ in layer_shapes():
        n = o * 9 * (a + b)
        ws.append(wflat[wo:wo + n].reshape(o, 9, a + b))
        bs.append(bflat[bo:bo + o])
        wo += n
        bo += o
    return ws, bs


def unet_loss_and_grads(feat, base, target, wflat, bflat, bf16_acts=True):
    """loss and gradients (flat, float64) of the whole network"""
    ws, bs = split_params(wflat, bflat)
    q = to_bf16 if bf16_acts else (lambda a: a)
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    A = {}

    def conv(i, x):
        return conv3x3(x, ws[i], bs[i])

    x = feat.astype(np.float64)
    A["a0"] = q(relu(conv(0, x)))
    A["e0"] = q(relu(conv(1, A["a0"])))
    A["p0"] = q(pool(A["e0"]))
    A["a1"] = q(relu(conv(2, A["p0"])))
    A["e1"] = q(relu(conv(3, A["a1"])))
    A["p1"] = q(pool(A["e1"]))
    A["a2"] = q(relu(conv(4, A["p1"])))
    A["e2"] = q(relu(conv(5, A["a2"])))
    A["p2"] = q(pool(A["e2"]))
    A["a3"] = q(relu(conv(6, A["p2"])))
    A["b"] = q(relu(conv(7, A["a3"])))
    A["u2"] = up(A["b"])
    A["cat2"] = np.concatenate([A["u2"], A["e2"]], axis=-1)
    A["c2"] = q(relu(conv(8, A["cat2"])))
    A["d2"] = q(relu(conv(9, A["c2"])))
    A["u1"] = up(A["d2"])
    A["cat1"] = np.concatenate([A["u1"], A["e1"]], axis=-1)
    A["c1"] = q(relu(conv(10, A["cat1"])))
    A["d1"] = q(relu(conv(11, A["c1"])))
    A["u0"] = up(A["d1"])
    A["cat0"] = np.concatenate([A["u0"], A["e0"]], axis=-1)
    A["c0"] = q(relu(conv(12, A["cat0"])))
    A["d0"] = q(relu(conv(13, A["c0"])))
    r = conv(14, A["d0"])
    out = base + r[..., :3]
    n = out.size
    loss = float(np.mean((out - target) ** 2))
    dr = np.zeros_like(r)
    dr[..., :3] = 2.0 * (out - target) / n

    gw = [None] * 15
    gb = [None] * 15

    def back(i, x, y, gy, act):
        gpre = gy * (y > 0) if act else gy
        gw[i] = conv3x3_wgrad(x, gpre)
        gb[i] = gpre.reshape(-1, gpre.shape[-1]).sum(axis=0)
        return conv3x3_dgrad(gpre, ws[i])

    g = back(14, A["d0"], None, dr, 0)
    g = back(13, A["c0"], A["d0"], g, 1)
    g = back(12, A["cat0"], A["c0"], g, 1)
    g_u0, g_e0 = g[..., :CH[1]], g[..., CH[1]:]
    g = back(11, A["c1"], A["d1"], up_bwd(g_u0), 1)
    g = back(10, A["cat1"], A["c1"], g, 1)
    g_u1, g_e1 = g[..., :CH[2]], g[..., CH[2]:]
    g = back(9, A["c2"], A["d2"], up_bwd(g_u1), 1)
    g = back(8, A["cat2"], A["c2"], g, 1)
    g_u2, g_e2 = g[..., :CH[3]], g[..., CH[3]:]
    g = back(7, A["a3"], A["b"], up_bwd(g_u2), 1)
    g = back(6, A["p2"], A["a3"], g, 1)
    g = back(5, A["a2"], A["e2"], pool_bwd(g) + g_e2, 1)
    g = back(4, A["p1"], A["a2"], g, 1)
    g = back(3, A["a1"], A["e1"], pool_bwd(g) + g_e1, 1)
    g = back(2, A["p0"], A["a1"], g, 1)
    g = back(1, A["a0"], A["e0"], pool_bwd(g) + g_e0, 1)
    back(0, x, A["a0"], g, 1)
    return loss, np.concatenate([v.ravel() for v in gw]), np.concatenate(gb), out


def unet_forward(feat, base, wflat, bflat):
    """inference output clamp(base + U-Net(feat)[..., :3], 0, 1) in float64
    with BF16-rounded activations (pooled values rounded too, as the fused
    GPU path rounds them when it reads them)"""
    ws, bs = split_params(wflat, bflat)
    q = to_bf16
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731

    def conv(i, x):
        return conv3x3(x, ws[i], bs[i])

    x = feat.astype(np.float64)
    a0 = q(relu(conv(0, x)))
    e0 = q(relu(conv(1, a0)))
    a1 = q(relu(conv(2, q(pool(e0)))))
    e1 = q(relu(conv(3, a1)))
    a2 = q(relu(conv(4, q(pool(e1)))))
    e2 = q(relu(conv(5, a2)))
    a3 = q(relu(conv(6, q(pool(e2)))))
    b = q(relu(conv(7, a3)))
    c2 = q(relu(conv(8, np.concatenate([up(b), e2], axis=-1))))
    d2 = q(relu(conv(9, c2)))
    c1 = q(relu(conv(10, np.concatenate([up(d2), e1], axis=-1))))
    d1 = q(relu(conv(11, c1)))
    c0 = q(relu(conv(12, np.concatenate([up(d1), e0], axis=-1))))
    d0 = q(relu(conv(13, c0)))
    r = conv(14, d0)
    return np.clip(base + r[..., :3], 0.0, 1.0)

"""Pins the oracle itself: published vectors, closed forms and invariants."""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import flame as fl
from refraktor_amd import render

# Random123 known-answer vectors for Philox4x32-10 (ctr, key, output)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_kat(ctr, key, out):
    assert tuple(oracle.c_philox(ctr, key)) == out
    assert tuple(oracle.philox(np.array([ctr]), *key)[0]) == out


def test_philox_numpy_matches_c():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, size=(64, 4), dtype=np.uint64).astype(np.uint32)
    key = (123456789, 987654321)
    v = oracle.philox(ctr, *key)
    for i in range(64):
        assert np.array_equal(v[i], oracle.c_philox(ctr[i], key))


def _identity_flame(weights=(1.0,), variation="linear"):
    xfs = [fl.xform(w, 0.0, 0.0, [fl.vlink([1, 0, 0, 1, 0, 0], {variation: 1.0})]) for w in weights]
    pal = [[0.0, 0.0, 1.0]] * 256
    return fl.make_flame(xfs, pal)


def _prepare(f, blocks=4, w=64, h=64, seed=1):
    return oracle.prepare(f, render.CONFIG_DIR, w, h, blocks, seed)


def test_permutation_is_a_bijection():
    """identity maps: one pass only permutes points within each block"""
    f = _identity_flame()
    L, P, pal = _prepare(f)
    n = 4 * 256
    x = np.arange(n, dtype=np.float64)
    y = -x
    c = np.zeros(n)
    tr, _ = oracle.trace(L, P, P.shape[1], 4, 256, 1, 0, 3, x, y, c)
    # trace holds each thread's output before the permutation; the outputs of
    # pass 2 must be a permutation of the block's inputs
    for b in range(4):
        sl = slice(b * 256, (b + 1) * 256)
        # polar interpolation leaves ~1e-16 in the off-diagonal affine terms
        np.testing.assert_allclose(np.sort(tr[2, sl, 0]), np.sort(x[sl]), atol=1e-9)
    assert not np.array_equal(tr[2, :, 0], x), "points should move between threads"


def test_selection_frequencies_follow_weights():
    w = (0.2, 0.3, 0.5)
    f = _identity_flame(w)
    L, P, pal = _prepare(f, blocks=64)
    n = 64 * 256
    z = np.zeros(n)
    _, xid = oracle.trace(L, P, P.shape[1], 64, 256, 5, 0, 200, z, z, z)
    counts = np.bincount(xid.ravel(), minlength=3) / xid.size
    np.testing.assert_allclose(counts, w, atol=0.01)


def test_zero_weight_xform_never_chosen():
    f = _identity_flame((0.5, 0.0, 0.5, 0.0))
    L, P, pal = _prepare(f, blocks=16)
    z = np.zeros(16 * 256)
    _, xid = oracle.trace(L, P, P.shape[1], 16, 256, 5, 0, 100, z, z, z)
    assert set(np.unique(xid)) <= {0, 2}


def test_last_xform_fallback():
    """when the last weight is zero, a ratio past the cumulative sum picks the last non-zero"""
    f = _identity_flame((1.0, 0.0))
    L, P, pal = _prepare(f, blocks=4)
    z = np.zeros(4 * 256)
    _, xid = oracle.trace(L, P, P.shape[1], 4, 256, 5, 0, 50, z, z, z)
    assert set(np.unique(xid)) == {0}


def test_hammersley_seeding_closed_form():
    f = _identity_flame()
    w = h = 64
    L, P, pal = _prepare(f, blocks=4, w=w, h=h)
    x, y, c = oracle.seed_walkers(P, P.shape[1], 4, 256, w, h, 3)
    n = 1024
    i = np.arange(n)
    rev = np.array([int(f"{v:010b}"[::-1], 2) for v in i])
    hx, hy = i / n * 2 - 1, rev / n * 2 - 1
    px, py = (hx + 1) / 2 * w, (hy + 1) / 2 * h
    ps = oracle.plane_space(f, w, h)
    np.testing.assert_allclose(x, ps[0] * px + ps[2] * py + ps[4], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(y, ps[1] * px + ps[3] * py + ps[5], rtol=1e-6, atol=1e-6)
    assert np.all((c >= 0) & (c <= 1))


def test_plane_and_screen_are_inverse():
    f = fl.flame_1080(2)
    f["rotate"] = 33.0
    f["center_x"], f["center_y"] = 0.3, -0.2
    s, p = oracle.screen_space(f, 1920, 1080), oracle.plane_space(f, 1920, 1080)
    pts = np.array([[0, 0], [1920, 1080], [100, 700]], dtype=np.float64)
    for px, py in pts:
        qx = p[0] * px + p[2] * py + p[4]
        qy = p[1] * px + p[3] * py + p[5]
        sx = s[0] * qx + s[2] * qy + s[4]
        sy = s[1] * qx + s[3] * qy + s[5]
        assert abs(sx - px) < 1e-9 and abs(sy - py) < 1e-9


def test_bin_points_closed_form():
    """hand-computed increments: palette lerp, rounding, bounds and the -0 column"""
    S = np.array([1, 0, 0, 1, 0, 0], dtype=np.float32)  # plane == pixel coords
    pal = np.arange(256, dtype=np.uint32) * 0x010101      # grey ramp, entry i = (i, i, i)
    pts = np.array([[3.5, 2.25, 0.0, 1.0],   # pixel (3,2): palette[0], W += 255
                    [3.9, 2.0, 1.0, 0.5],    # pixel (3,2): 255*0.5 -> 128 per channel, W += 128
                    [-0.5, 0.0, 0.5, 1.0],   # trunc(-0.5) = -0 -> column 0; lerp(127,128,0.5)=127.5 -> 128
                    [10.0, 0.0, 0.0, 1.0],   # x = width: outside
                    [np.nan, 0, 0, 1],       # non-finite: no draw
                    [0.0, 0.0, 0.0, 0.0]],   # zero opacity: no draw
                   dtype=np.float32)
    hst = oracle.bin_points(S, pal, pts, 10, 5)
    assert tuple(hst[2, 3]) == (128, 128, 128, 383)
    assert tuple(hst[0, 0]) == (128, 128, 128, 255)
    assert int(hst.astype(np.int64).sum()) == 128 * 3 + 383 + 128 * 3 + 255


def test_catmull_invariants():
    ss = 5
    # constant samples -> constant segment
    s = np.tile(np.array([1.5, -2.0, 0.0, 3.0, 7.0]), 4)
    seg = oracle.catmull(s, ss)
    assert np.allclose(seg[0, :, :3], 0) and np.allclose(seg[0, :, 3], s[:ss])
    # equally spaced collinear samples -> linear in t
    s = np.concatenate([np.full(ss, v) for v in (0.0, 1.0, 2.0, 3.0)])
    seg = oracle.catmull(s, ss)[0]
    for t in (0.0, 0.25, 0.5, 1.0):
        val = seg[:, 0] * t ** 3 + seg[:, 1] * t ** 2 + seg[:, 2] * t + seg[:, 3]
        np.testing.assert_allclose(val, 1.0 + t, atol=1e-12)
    # interpolating: p1 at t = 0 and p2 at t = 1
    rng = np.random.default_rng(2)
    s = rng.normal(size=4 * ss)
    seg = oracle.catmull(s, ss)[0]
    np.testing.assert_allclose(seg[:, 3], s[ss:2 * ss])
    np.testing.assert_allclose(seg.sum(axis=1), s[2 * ss:3 * ss], atol=1e-12)


def test_tonemap_special_cases():
    hst = np.zeros((1, 3, 4), dtype=np.uint32)
    hst[0, 1] = (255 * 40, 255 * 20, 255 * 10, 255 * 50)
    hst[0, 2] = (255 * 3, 255 * 3, 255 * 3, 255 * 3)
    q, g, b = 2.0, 2.2, 4.0
    out0 = oracle.tonemap(hst, q, g, b, 0.0)
    out1 = oracle.tonemap(hst, q, g, b, 1.0)
    assert np.all(out0[0, 0] == 0) and np.all(out1[0, 0] == 0)
    k = 10 ** (-math.log10(q) - 0.5)
    w = 50.0
    fac = 0.5 * b * math.log10(1 + w * k) / w
    c = np.array([40.0, 20.0, 10.0]) * fac
    np.testing.assert_allclose(out0[0, 1, :3], np.clip(c ** (1 / g), 0, 1))          # vibrancy 0: c^(1/gamma)
    np.testing.assert_allclose(out1[0, 1, :3], np.clip(c * (w * fac) ** (1 / g - 1), 0, 1))  # vibrancy 1
    assert abs(out1[0, 1, 3] - math.log2(50.0) / 24.0) < 1e-12


def test_density_estimate_radius_zero_is_identity():
    rng = np.random.default_rng(4)
    hst = rng.integers(0, 2000, size=(12, 15, 4)).astype(np.uint32)
    hst[rng.random((12, 15)) < 0.3] = 0
    out = oracle.density_estimate(hst, 0.0, 0.0, 0.4)
    np.testing.assert_allclose(out, hst / 255.0, atol=1e-12)


def test_density_estimate_conserves_energy():
    hst = np.zeros((40, 40, 4), dtype=np.uint32)
    hst[20, 20] = (255, 510, 765, 255)       # density 1 -> widest kernel
    hst[10, 25] = (2550, 0, 0, 255 * 1000)   # dense bin -> narrow kernel
    out = oracle.density_estimate(hst, 6.0, 0.0, 0.4)
    np.testing.assert_allclose(out.sum(axis=(0, 1)), hst.sum(axis=(0, 1)) / 255.0, rtol=1e-12)
    # the spread of a density-1 bin is symmetric and peaks at its centre
    blob = out[14:27, 14:27, 3]
    assert np.allclose(blob, blob[::-1, ::-1]) and blob.argmax() == 6 * 13 + 6


def test_oracle_render_counts():
    f = fl.flame_small(2)
    L, P, pal = _prepare(f, blocks=8, w=80, h=45)
    hst, draws = oracle.render(L, P, P.shape[1], pal, 8, 256, 80, 45, 2, 5, 32)
    assert draws <= 8 * 256 * 32
    assert int(hst[..., 3].astype(np.int64).sum()) == draws * 255  # opacity 1 everywhere


def test_density_hdr_closed_forms():
    """gain (cdf/0.5)^2 on the HSV value; hue and saturation kept (colorsys)"""
    import colorsys
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (64, 4))
    bins = np.full(97, 10)  # uniform CDF: cdf(k) = (k + 1) / 97
    out = oracle.density_hdr(img[None], bins)[0]
    for px, o in zip(img, out):
        power = px[3] * 96 + 1
        lo = min(int(power), 96)
        hi = min(lo + 1, 96)
        v = ((lo + 1) + (power - np.floor(power)) * ((hi + 1) - (lo + 1))) / 97
        h, s, val = colorsys.rgb_to_hsv(*px[:3])
        want = colorsys.hsv_to_rgb(h, s, val * (v / 0.5) ** 2)
        np.testing.assert_allclose(o[:3], want, rtol=1e-5, atol=1e-6)
        assert o[3] == 1.0
    # at cdf = 0.5 the image is unchanged; gray stays gray
    bins = np.zeros(97)
    bins[0] = bins[96] = 1  # cdf: 0.5 up to bin 95, 1 at bin 96
    px = np.array([[[0.3, 0.3, 0.3, 10 / 96.0], [0.2, 0.5, 0.9, 20 / 96.0]]])
    np.testing.assert_allclose(oracle.density_hdr(px, bins)[0, :, :3], px[0, :, :3], atol=1e-7)


def test_convert_image_closed_forms():
    """8-bit channels truncate min(v * 255, 255) (convert.cu:46-48): 0.5 ->
    127, not 128; half(1/255) sits just below 1/255 and gives 0; values above
    1 clamp to 255; a half3 input gets alpha 1; swap exchanges R and B; flip
    reverses the rows"""
    x = np.array([[[1.0, 0.5, 0.0], [2.0, 0.25, 1 / 255]],
                  [[0.75, 0.125, 0.999], [0.0, 1.0, 0.5]]], np.float16)
    rgba = oracle.convert_image(x, 4)
    assert rgba.dtype == np.uint8 and rgba.shape == (2, 2, 4)
    assert rgba[0, 0].tolist() == [255, 127, 0, 255]
    assert rgba[0, 1].tolist() == [255, 63, 0, 255]
    assert float(np.float16(1 / 255)) * 255 < 1.0
    assert rgba[1, 0].tolist() == [191, 31, int(float(np.float16(0.999)) * 255), 255]
    rgb_sw = oracle.convert_image(x, 5, swap=True)
    assert rgb_sw.shape == (2, 2, 3) and rgb_sw[0, 0].tolist() == [0, 127, 255]
    fl4 = oracle.convert_image(x, 0, flip=True)
    assert fl4.dtype == np.float32 and fl4[0, 1].tolist() == [0.0, 1.0, 0.5, 1.0]
    np.testing.assert_array_equal(oracle.convert_image(x, 2), x.astype(np.float32))
    h4 = np.concatenate([x, np.full((2, 2, 1), 0.5, np.float16)], axis=-1)
    assert oracle.convert_image(h4, 4)[..., 3].tolist() == [[127, 127], [127, 127]]
    np.testing.assert_array_equal(oracle.convert_image(h4, 3, swap=True, flip=True), h4[::-1, :, [2, 1, 0]])


def _jpeg_images():
    y, x = np.mgrid[0:45, 0:70]
    smooth = np.stack([x * 255 // 70, y * 255 // 45, (x + y) * 255 // 115], -1).astype(np.uint8)
    noise = np.random.default_rng(3).integers(0, 256, size=(61, 83, 3), dtype=np.uint8)  # 11 x 8 MCUs
    return smooth, noise


def _decode(data: bytes) -> np.ndarray:
    import io
    from PIL import Image
    im = Image.open(io.BytesIO(data))
    assert im.format == "JPEG"
    return np.asarray(im.convert("RGB")).astype(np.float64)


def _psnr(a, b) -> float:
    mse = float(((np.asarray(a, np.float64) - b) ** 2).mean())
    return 99.0 if mse == 0 else 10 * math.log10(255.0 ** 2 / mse)


@pytest.mark.parametrize("quality", [1, 10, 50, 75, 90, 100])
def test_jpeg_oracle_decodes_like_libjpeg(quality):
    """An independent decoder (PIL's libjpeg) reads the oracle's files, and
    the decoded image is as close to the input as libjpeg's own 4:4:4
    encoding at the same quality (same standard tables and IJG scaling):
    within 0.5 dB and 10 % of its size."""
    import io
    from PIL import Image
    smooth, _ = _jpeg_images()
    data = oracle.jpeg_encode(smooth, quality)
    assert data[:2] == b"\xff\xd8" and data[-2:] == b"\xff\xd9"
    ours = _psnr(_decode(data), smooth)
    buf = io.BytesIO()
    Image.fromarray(smooth).save(buf, "JPEG", quality=quality, subsampling=0)
    ref = _psnr(_decode(buf.getvalue()), smooth)
    assert abs(ours - ref) < 0.5, (ours, ref)
    assert abs(len(data) - len(buf.getvalue())) < 0.1 * len(buf.getvalue())


def test_jpeg_oracle_stress_cases():
    """noise at quality 100 (10-bit AC magnitudes, many 0xFF bytes stuffed),
    RGBA input, sizes that are not whole blocks, a single pixel, and enough
    MCUs for every restart marker RST0..RST7 to appear"""
    _, noise = _jpeg_images()
    data = oracle.jpeg_encode(noise, 100)
    assert _psnr(_decode(data), noise) > 40.0
    body = data[data.index(b"\xff\xda"):]
    assert all(bytes([0xFF, 0xD0 + n]) in body for n in range(8))
    rgba = np.concatenate([noise, np.full(noise.shape[:2] + (1,), 7, np.uint8)], -1)
    assert oracle.jpeg_encode(rgba, 100) == data  # alpha is ignored
    one = np.array([[[200, 30, 90]]], np.uint8)
    np.testing.assert_allclose(_decode(oracle.jpeg_encode(one, 95))[0, 0], one[0, 0], atol=3)
    gray = np.full((9, 17, 3), 128, np.uint8)  # mid grey: every coefficient zero
    np.testing.assert_array_equal(_decode(oracle.jpeg_encode(gray, 50)), gray)


# ------------------------------------------------ independent pins (restated
# functions checked against textbook forms or hand-computed values)

def test_hsv_to_rgb8_matches_colorsys():
    """refactor.cu:119-151 f(n) = v - v s max(0, min(k, 4 - k, 1)) with
    k = fmod(n + h / 60, 6) is the textbook HSV -> RGB (colorsys) for hues in
    [0, 360); bytes are trunc(255 c), R in the low byte, alpha 255"""
    import colorsys
    rng = np.random.default_rng(4)
    h, s, v = rng.uniform(0, 360, 2000), rng.uniform(0, 1, 2000), rng.uniform(0, 1, 2000)
    packed = oracle.hsv_to_rgb8(h, s, v)
    got = np.stack([(packed >> sh) & 0xFF for sh in (0, 8, 16, 24)], axis=1)
    want = np.array([colorsys.hsv_to_rgb(a / 360.0, b, c) for a, b, c in zip(h, s, v)]) * 255.0
    near_int = np.abs(want - np.round(want)) < 1e-6
    d = got[:, :3].astype(np.int64) - np.floor(want).astype(np.int64)
    assert np.all((d == 0) | (near_int & (np.abs(d) <= 1)))
    assert np.all(got[:, 3] == 255)
    # primaries, grey, and a negative hue (C fmod keeps the sign: -60 wraps
    # to magenta here, but below -300 a channel's k goes negative and the
    # channel stays at v, as the reference computes it)
    cases = {(0.0, 1.0, 1.0): (255, 0, 0), (120.0, 1.0, 1.0): (0, 255, 0), (240.0, 1.0, 1.0): (0, 0, 255),
             (77.0, 0.0, 0.5): (127, 127, 127), (-60.0, 1.0, 1.0): (255, 0, 255), (-400.0, 1.0, 1.0): (255, 255, 255)}
    for (hh, ss, vv), rgb in cases.items():
        p = int(oracle.hsv_to_rgb8(np.array([hh]), np.array([ss]), np.array([vv]))[0])
        assert (p & 0xFF, (p >> 8) & 0xFF, (p >> 16) & 0xFF) == rgb, (hh, ss, vv)


def test_density_histogram_hand_computed():
    """density_histo.cu:1-41: bin = trunc(ln(v) * 4 / ln 2) + 1 for a density
    v >= 1 draw (v = W / 255), else bin 0; max = trunc(max v)"""
    W = np.array([0, 127, 255, 383, 765, 255000], dtype=np.uint32)
    # v: 0, 0.498, 1, 1.502 (log2 0.587 -> 2.35), 3 (log2 1.585 -> 6.34), 1000 (log2 9.966 -> 39.86)
    hist = np.zeros((1, W.size, 4), np.uint32)
    hist[0, :, 3] = W
    bins, mx = oracle.density_histogram(hist)
    want = np.zeros(97, np.uint64)
    for b, n in {0: 2, 1: 1, 3: 1, 7: 1, 40: 1}.items():
        want[b] = n
    assert np.array_equal(bins, want)
    assert mx == 1000


def test_precalc_lambdas_match_the_database_programs():
    """the oracle's precalc of waves, julian and perspective (used by its
    warmup) against the database's precalc programs run by the interpreter,
    which the golden vectors pin to librefrakt's"""
    from oracle import flang_eval as fe
    from refraktor_amd import render
    cfg = Path(render.CONFIG_DIR)
    mdb, _ = fe.load_db((cfg / "variations_new.yml").read_text(), (cfg / "common.yml").read_text())
    rng = np.random.default_rng(8)
    for trial in range(20):
        aff = rng.uniform(-1.5, 1.5, 6)
        user = {"power": float(rng.choice([-3.0, -2.0, 2.0, 5.0]) * rng.uniform(0.5, 1.5)),
                "dist": float(rng.uniform(-2, 2)), "angle": float(rng.uniform(-3, 3))}
        f = fl.make_flame([fl.xform(1.0, 0.0, 0.5, [fl.vlink(list(aff), {
            "waves": 1.0,
            "julian": {"weight": 1.0, "parameters": {"power": user["power"], "dist": user["dist"]}},
            "perspective": {"weight": 1.0, "parameters": {"angle": user["angle"], "dist": user["dist"]}}})])],
            [[0.0, 0.0, 1.0]] * 256)
        L = oracle.Layout(f, oracle.load_variation_params(cfg))
        P = np.zeros((1, L.size_reals + 8))
        vs = L.xforms[0]["vlinks"][0]
        P[0, vs["aff"]:vs["aff"] + 6] = aff
        for slot in vs["vars"]:
            for pn in slot["params"]:
                if pn in user:
                    P[0, slot["params"][pn]] = user[pn]
        got = oracle.precalc(P, L, 1, 0)[0]
        a, d, b, e, c, ff = aff  # (a, d, b, e, c, f) order in the layout
        affd = {"a": a, "b": b, "c": c, "d": d, "e": e, "f": ff}
        for slot in vs["vars"]:
            v = mdb[slot["name"]]
            want = fe.precalc(v, {pn: P[0, slot["params"][pn]] for pn in v.user_params}, affd)
            for pn, off in slot["params"].items():
                assert abs(got[off] - want[pn]) <= 1e-12 * max(1.0, abs(want[pn])), (slot["name"], pn)

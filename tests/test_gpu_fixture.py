"""End to end on the reference's own test flame
(assets/flames_test/electricsheep.247.50049.flam3, copied as data into
tests/golden): flam3 import, GPU render against the oracle's render of the
same streams, queue binning bit-exact against atomic binning, density
estimation, tonemap and the 8-bit conversion."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from parity_stats import TAU, converged_ratio
from refraktor_amd import _lib
from refraktor_amd import render

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "electricsheep.247.50049.flam3"
W, H = 200, 148


def test_flam3_fixture_end_to_end(gpu, db):
    fo = db.flame_from_flam3(GOLDEN.read_text())
    js = fo.to_json()
    # 2^26 iterations: 2^16 walkers x 1024 passes
    walkers, passes = 1 << 16, 1024
    k = db.compile(fo)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=21, fuse=20, walkers=walkers)
    res = st.bin(iters=walkers * passes)
    g = st.histogram()
    inf = st.info()
    assert inf.quality > 0 and int(g[..., 3].astype(np.uint64).sum()) == res["hit_sum"]

    # per-pixel relative L1 against the oracle's render from independent
    # streams, within TAU of the oracle's own seed-to-seed figure
    L = oracle.Layout(js, oracle.load_variation_params(render.CONFIG_DIR))
    ref, _ = oracle.render(L, st.params(), inf.param_stride, st.palettes(), inf.blocks, inf.threads_per_block, W, H,
                           22, 20, passes)
    other, _ = oracle.render(L, st.params(), inf.param_stride, st.palettes(), inf.blocks, inf.threads_per_block, W,
                             H, 23, 20, passes)
    ratio, s_gpu, s_seed, n = converged_ratio(g, ref, other)
    print(f"[flam3 fixture] pixels {n} gpu {s_gpu:.4f} oracle seeds {s_seed:.4f} ratio {ratio:.3f}")
    assert n > 2000
    assert ratio <= TAU, (ratio, s_gpu, s_seed)

    kq = db.compile(fo, bin_queue=True)
    sq = kq.warmup(kq.pack_samples(W, H), W, H, seed=21, fuse=20, walkers=walkers)
    sq.bin(iters=walkers * passes)
    assert np.array_equal(sq.histogram(), g)

    de = st.density_estimate(6.0)
    assert np.isfinite(de).all() and de.min() >= 0.0 and de[..., 3].sum() > 0.0
    half = st.tonemap(js.get("gamma", 4.0), js.get("brightness", 4.0), js.get("vibrancy", 1.0), fmt=_lib.PIXEL_HALF4)
    rgba = render.convert_image(half, _lib.PIXEL_RGBA8, flip=True)
    np.testing.assert_array_equal(rgba, oracle.convert_image(half, _lib.PIXEL_RGBA8, flip=True))
    assert (rgba[..., :3] > 0).any()
    assert not ((rgba[::-1][..., 3] > 0) & (g[..., 3] == 0)).any()  # alpha only where draws landed

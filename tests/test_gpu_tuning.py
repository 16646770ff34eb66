"""Launch shapes: every shape the autotuner may pick renders the same
histogram (integer binning is order-free), and the tuner picks a shape."""
import numpy as np
import pytest

from refraktor_amd import autotune
from refraktor_amd import flame as fl

pytestmark = pytest.mark.gpu

W, H = 320, 180


def _render(db, flm, tpb, aw, tile, seed=4, hist_bits=32, queue=False, walkers=1 << 14, iters=1 << 21):
    k = db.compile(flm, threads_per_block=tpb, atomic_width=aw, bin_tile_log2=tile, hist_bits=hist_bits,
                   bin_queue=queue)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=seed, walkers=walkers)
    res = st.bin(iters=iters)
    return st, res


@pytest.mark.parametrize("name", ["small", "1080"])
def test_atomic_width_and_bin_tiles_bit_exact(gpu, db, name):
    flm = fl.flame_small(2) if name == "small" else fl.flame_1080(2)
    rst, rres = _render(db, flm, 256, 32, 0)
    ref = rst.histogram().astype(np.uint64)
    assert rres["total_draws"] > 0
    for aw, tile, hb in ((64, 0, 32), (32, 10, 32), (64, 10, 32), (32, 6, 32), (32, 12, 32), (64, 12, 32),
                         (32, 0, 64), (32, 11, 64)):
        st, res = _render(db, flm, 256, aw, tile, hist_bits=hb)
        assert res["total_draws"] == rres["total_draws"]
        got = st.histogram()
        assert got.dtype == (np.uint64 if hb == 64 else np.uint32)
        np.testing.assert_array_equal(got.astype(np.uint64), ref, err_msg=f"atomic {aw} tile {tile} bins {hb}")


@pytest.mark.parametrize("name", ["small", "1080"])
def test_queue_binning_bit_exact(gpu, db, name):
    """records sorted by screen tile and summed in LDS give the atomic histogram"""
    flm = fl.flame_small(3) if name == "small" else fl.flame_1080(3)
    refs = {}
    for tpb in (256, 1024):  # walkers per block change the streams: one reference each
        rst, rres = _render(db, flm, tpb, 32, 0)
        refs[tpb] = (rst.histogram().astype(np.uint64), rres)
    for aw, hb, tpb in ((32, 32, 256), (64, 32, 256), (32, 64, 1024)):
        ref, rres = refs[tpb]
        st, res = _render(db, flm, tpb, aw, 0, hist_bits=hb, queue=True)
        assert res["total_draws"] == rres["total_draws"] and res["total_passes"] == rres["total_passes"]
        np.testing.assert_array_equal(st.histogram().astype(np.uint64), ref, err_msg=f"queue aw {aw} bins {hb}")


def test_queue_binning_many_chunks(gpu, db):
    """2^20 walkers x 200 passes: four record chunks of 64 passes and a short one"""
    flm = fl.flame_1080(4)
    args = dict(walkers=1 << 20, iters=200 << 20)
    a, ra = _render(db, flm, 1024, 64, 12, **args)
    b, rb = _render(db, flm, 1024, 64, 0, queue=True, **args)
    assert ra["total_passes"] == rb["total_passes"] == 200 << 20
    np.testing.assert_array_equal(a.histogram(), b.histogram())


def test_queue_binning_bailouts(gpu, db):
    k = db.compile(fl.flame_1080(2), threads_per_block=1024, atomic_width=64, bin_queue=True)
    s = k.pack_samples(W, H)
    # quality: stops once the whole call reached the target mean density
    st = k.warmup(s, W, H, seed=3, walkers=1 << 16)
    r = st.bin(quality=20.0)
    assert r["quality"] >= 20.0 and r["total_passes"] % 1024 == 0  # blocks stop at their own checks
    assert r["quality"] < 60.0  # stopped within a few bailout checks
    # time: a few milliseconds of device clock
    st = k.warmup(s, W, H, seed=3, walkers=1 << 16)
    r = st.bin(millis=5.0)
    assert 0 < r["total_passes"] and r["kernel_ms"] < 500.0
    # the histogram holds exactly the counted draws
    h = st.histogram()
    assert int(h[..., 3].sum(dtype=np.uint64)) == r["hit_sum"]


def test_64bit_bins_feed_the_same_image(gpu, db):
    """density histogram, density estimation and tonemap read u64 bins as they read u32 bins"""
    flm = fl.flame_1080(5)
    a, _ = _render(db, flm, 256, 32, 0)
    b, _ = _render(db, flm, 256, 32, 0, hist_bits=64)
    ba, ma = a.density_histogram()
    bb, mb = b.density_histogram()
    np.testing.assert_array_equal(ba, bb)
    assert ma == mb
    np.testing.assert_array_equal(a.density_estimate(6.0), b.density_estimate(6.0))
    np.testing.assert_array_equal(a.tonemap(2.2, 4.0, 1.0), b.tonemap(2.2, 4.0, 1.0))


def test_64bit_bins_do_not_wrap(gpu, db):
    """one pixel taking more than 2^32 / 255 full-opacity draws: u64 bins hold the count"""
    k = db.compile(fl.flame_small(1), hist_bits=64)
    st = k.warmup(k.pack_samples(64, 64), 64, 64, seed=1, walkers=1 << 12)
    S = st.params()[0, :6]
    # plane point that lands in pixel (10, 20): solve the screen affine
    A = np.array([[S[0], S[2]], [S[1], S[3]]], dtype=np.float64)
    x, y = np.linalg.solve(A, np.array([20.5 - S[4], 10.5 - S[5]]))
    n = 17_000_000  # 255 * n > 2^32
    pts = np.tile(np.array([[x, y, 0.5, 1.0]], dtype=np.float32), (n, 1))
    st.clear_histogram()
    st.bin_points(pts)
    h = st.histogram()
    assert int(h[10, 20, 3]) == 255 * n > (1 << 32)
    assert int(h.sum(dtype=np.uint64)) == int(h[10, 20].sum(dtype=np.uint64))


def test_bad_tile_sizes_rejected(gpu, db):
    with pytest.raises(Exception):
        db.compile(fl.flame_small(1), threads_per_block=1024, bin_tile_log2=13)


def test_autotune_picks_a_running_shape(gpu, db, tmp_path, monkeypatch):
    monkeypatch.setenv("RFK_CACHE_DIR", str(tmp_path))
    f = db.flame(fl.flame_small(3))
    samples = f.pack_samples(W, H)
    grid = [autotune.Shape(256, 32, 0, 1 << 16), autotune.Shape(512, 64, 10, 1 << 16),
            autotune.Shape(256, 32, 13, 1 << 16)]  # the last is not a valid tile size
    shape, results = autotune.tune(db, f, samples, W, H, grid=grid, probe_iters=1 << 22)
    assert shape in grid[:2]
    assert results[2]["status"] != 0 and results[2]["iters_per_s"] == 0.0
    assert all(r["iters_per_s"] > 0 for r in results[:2])
    again, _ = autotune.tune(db, f, samples, W, H, grid=grid, probe_iters=1 << 22)
    assert again == shape  # served from the cache file


def test_animation_autotuned_batch(gpu, db):
    """the animation picks its launch shape with the autotuner (one frame of
    the aligned keyframes at the batch's size and walkers per frame) and the
    batch renders every frame at its budget with that shape"""
    from refraktor_amd import animation
    kf = animation.keyframe_flames(4, 3)
    walkers = 1 << 14
    shape, results = animation.tune_shape(db, kf, W, H, walkers, probe_iters=1 << 21, use_cache=False)
    assert shape in animation.batch_grid(walkers)
    assert all(r["status"] == 0 and r["iters_per_s"] > 0 for r in results)
    assert sum(r["best"] for r in results) == 1
    budgets = np.array([1 << 20, 3 << 18, 1 << 19, 5 << 18, 1 << 20, 1 << 18], dtype=np.uint64)
    anim = animation.Animation(db, kf, W, H, len(budgets), budgets, walkers=walkers, seed=3,
                               threads_per_block=shape.threads_per_block, atomic_width=shape.atomic_width,
                               bin_tile_log2=shape.bin_tile_log2, bin_queue=bool(shape.bin_queue))
    anim.render(graph=True)
    for f in range(len(budgets)):
        inf = anim.batch.frame(f)
        assert int(budgets[f]) <= inf["total_passes"] < int(budgets[f]) + inf["walkers"]
        h = anim.batch.histogram(f)
        assert int(h[..., 3].sum(dtype=np.uint64)) == 255 * inf["total_draws"]

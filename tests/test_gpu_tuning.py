"""Launch shapes: every shape the autotuner may pick renders the same
histogram (integer binning is order-free), and the tuner picks a shape."""
import numpy as np
import pytest

from refraktor_amd import autotune
from refraktor_amd import flame as fl

pytestmark = pytest.mark.gpu

W, H = 320, 180


def _render(db, flm, tpb, aw, tile, seed=4):
    k = db.compile(flm, threads_per_block=tpb, atomic_width=aw, bin_tile_log2=tile)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=seed, walkers=1 << 14)
    res = st.bin(iters=1 << 21)
    return st.histogram(), res


@pytest.mark.parametrize("name", ["small", "1080"])
def test_atomic_width_and_bin_tiles_bit_exact(gpu, db, name):
    flm = fl.flame_small(2) if name == "small" else fl.flame_1080(2)
    ref, rres = _render(db, flm, 256, 32, 0)
    assert rres["total_draws"] > 0
    for aw, tile in ((64, 0), (32, 10), (64, 10), (32, 6), (32, 12), (64, 12)):
        got, res = _render(db, flm, 256, aw, tile)
        assert res["total_draws"] == rres["total_draws"]
        np.testing.assert_array_equal(got, ref, err_msg=f"atomic {aw} tile {tile}")


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

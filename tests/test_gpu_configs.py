"""The 4K plain render (3840x2160, 12 xforms of up to 4 variations + final
xform, 2^32 iterations, density estimation radius 0-12) and the batched
animation (4 keyframes of 6-10 xforms Catmull-Rom interpolated into 120
frames at 1280x720, budgets log-uniform in [2^26, 2^29], queue-binned chunks
replayed from a hipGraph), each at its real size."""
import numpy as np
import pytest

import xform_bound as xb
from oracle import oracle
from parity_stats import TAU, converged_ratio
from refraktor_amd import animation
from refraktor_amd import flame as fl
from refraktor_amd import render

pytestmark = pytest.mark.gpu

W4K, H4K = 3840, 2160


def _layout(f):
    return oracle.Layout(f, oracle.load_variation_params(render.CONFIG_DIR))


# ------------------------------------------------------------------ 4K

def test_4k_flame_trajectory_one_step(db, gpu):
    """the 12-xform flame's steps within the stated bound (tests/xform_bound.py)"""
    f = fl.flame_4k(1)
    k = db.compile(f)
    st = k.warmup(k.pack_samples(320, 180), 320, 180, seed=5, fuse=3, walkers=4096)
    inf = st.info()
    x0 = st.walkers()
    tr, xid = st.trace(12)
    otr, oxid = oracle.trace(_layout(f), st.params(), inf.param_stride, inf.blocks, inf.threads_per_block, 5, 3, 12,
                             x0[:, 0], x0[:, 1], x0[:, 2], gpu_trace=tr)
    assert np.array_equal(xid, oxid)
    mdb, cm = xb.load_db(render.CONFIG_DIR)
    r = xb.trace_bound(mdb, cm, _layout(f), st.params(), inf.threads_per_block, 5, 0, 3, x0, tr, xid, 0, 8)
    print(f"[4k bound] {r}")
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["knife_steps"] <= 0.01 * r["steps"], r


@pytest.fixture(scope="module")
def k4(db, gpu):
    f = fl.flame_4k(1)
    return f, db.compile(f, threads_per_block=512, atomic_width=64, hist_bits=64, bin_queue=True)


def test_4k_full_budget(k4):
    """2^32 iterations into 64-bit bins: the exact budget, every draw counted
    once (sum of W = 255 x draws = hit sum, opacities are 1), no wrap"""
    f, k = k4
    walkers = 1 << 20
    st = k.warmup(k.pack_samples(W4K, H4K), W4K, H4K, seed=1, fuse=20, walkers=walkers)
    res = st.bin(iters=1 << 32)
    assert res["total_passes"] == 1 << 32
    h = st.histogram()
    assert h.dtype == np.uint64
    wsum = int(h[..., 3].sum(dtype=np.uint64))
    assert wsum == res["hit_sum"] == 255 * res["total_draws"]
    assert 0.5 * (1 << 32) < res["total_draws"] <= 1 << 32
    # colour channels carry at most 255 per draw
    assert int(h[..., :3].max()) <= int(h[..., 3].max())
    print(f"[4k] {res['total_draws']} draws, {res['kernel_ms']:.1f} ms, {int((h[..., 3] > 0).sum())} pixels hit")


@pytest.fixture(scope="module")
def k4_2_28(db, gpu, k4):
    """the queue-binned and the atomic 4K renders of 2^28 iterations"""
    f, kq = k4
    ka = db.compile(f, threads_per_block=512, atomic_width=64, hist_bits=64)
    out = []
    for k in (kq, ka):
        st = k.warmup(k.pack_samples(W4K, H4K), W4K, H4K, seed=2, fuse=20, walkers=1 << 20)
        res = st.bin(iters=1 << 28)
        assert res["total_passes"] == 1 << 28
        out.append(st)
    return out


def test_4k_queue_binning_equals_atomic(k4_2_28):
    sq, sa = k4_2_28
    hq, ha = sq.histogram(), sa.histogram()
    assert int(hq[..., 3].sum(dtype=np.uint64)) > 0
    assert np.array_equal(hq, ha)


def test_4k_density_estimate_radius_12(k4_2_28):
    """radius 0-12 (the 4K configuration) against the brute-force scatter on a
    64x64 crop, computed from the crop's sources and a margin of 12"""
    st = k4_2_28[0]
    h = st.histogram()
    got = st.density_estimate(12.0, 0.0, 0.4).astype(np.float64)
    w = h[..., 3].astype(np.float64)
    ys, xs = np.nonzero(w)
    cy = int(np.clip(np.average(ys, weights=w[ys, xs]), 80, H4K - 80))
    cx = int(np.clip(np.average(xs, weights=w[ys, xs]), 80, W4K - 80))
    R, S = 12, 64
    y0, x0 = cy - S // 2, cx - S // 2
    region = h[y0 - R:y0 + S + R, x0 - R:x0 + S + R]
    want = oracle.density_estimate(region, 12.0, 0.0, 0.4)[R:R + S, R:R + S]
    crop = got[y0:y0 + S, x0:x0 + S]
    assert (region[..., 3] > 0).mean() > 0.5
    np.testing.assert_allclose(crop, want, rtol=2e-4, atol=1e-4 * want.max())


# ------------------------------------------------------------------ animation

KF_SEED = 1
AW, AH, FRAMES = 1280, 720, 120


@pytest.fixture(scope="module")
def keyframes():
    return animation.keyframe_flames(4, KF_SEED)


def test_animation_full_batch(db, gpu, keyframes):
    """all 120 frames at their budgets, queue-binned and replayed from a
    hipGraph: every frame gets exactly its budget rounded up to whole passes,
    counts every draw once (sum of W = 255 x draws), and a replay reproduces
    the eager run bit for bit"""
    budgets = animation.frame_budgets(FRAMES, KF_SEED)
    assert budgets.min() >= 1 << 26 and budgets.max() <= 1 << 29
    anim = animation.Animation(db, keyframes, AW, AH, FRAMES, budgets, walkers=1 << 18, seed=KF_SEED,
                               threads_per_block=512, atomic_width=64, bin_queue=True)
    anim.render(graph=False)
    probe = {f: anim.batch.histogram(f) for f in (0, 37, 60, 119)}
    ms = anim.render(graph=True)
    print(f"[animation] 120 frames in {ms:.1f} ms (graph replay)")
    for f in range(FRAMES):
        inf = anim.batch.frame(f)
        walkers = inf["walkers"]
        assert inf["total_passes"] == inf["passes_per_walker"] * walkers
        assert int(budgets[f]) <= inf["total_passes"] < int(budgets[f]) + walkers
        h = anim.batch.histogram(f)
        wsum = int(h[..., 3].sum(dtype=np.uint64))
        assert wsum == 255 * inf["total_draws"], f
        assert 0 < inf["total_draws"] <= inf["total_passes"]
        if f in probe:
            assert np.array_equal(h, probe[f]), f


@pytest.mark.parametrize("frame", [0, 60, 119])
def test_animation_frames_match_oracle(db, gpu, keyframes, frame):
    """frames 0, 60 and 119 of the 120-frame batch at a reduced budget (2^25
    iterations each) against the oracle's render of the same frame's
    parameter rows from independent streams (tests/parity_stats.py)"""
    anim = _reduced_batch(db, keyframes)
    inf = anim.batch.frame(frame)
    g = anim.batch.histogram(frame)
    P, pal = anim.batch.params(frame), anim.batch.palettes(frame)
    L = _layout(anim.aligned[0].to_json())
    tpb = anim.kernel.threads_per_block

    def orender(seed):
        h, _ = oracle.render(L, P, inf["param_stride"], pal, inf["blocks"], tpb, AW, AH, seed, inf["fuse"],
                             inf["passes_per_walker"], stream=frame)
        return h

    ratio, s_gpu, s_seed, n = converged_ratio(g, orender(KF_SEED + 100), orender(KF_SEED + 200))
    print(f"[animation frame {frame}] pixels {n} gpu {s_gpu:.4f} oracle seeds {s_seed:.4f} ratio {ratio:.3f}")
    assert n > 2000
    assert ratio <= TAU, (ratio, s_gpu, s_seed)


_REDUCED = {}


def _reduced_batch(db, keyframes):
    if "anim" not in _REDUCED:
        anim = animation.Animation(db, keyframes, AW, AH, FRAMES, np.full(FRAMES, 1 << 25, np.uint64),
                                   walkers=1 << 18, seed=KF_SEED, threads_per_block=512, atomic_width=64,
                                   bin_queue=True)
        anim.render(graph=False)
        _REDUCED["anim"] = anim
    return _REDUCED["anim"]

"""GPU parity tests of the iterate/bin path against the oracle (through the C ABI)."""
import numpy as np
import pytest

from oracle import oracle
from parity_stats import TAU, converged_ratio
import xform_bound as xb
from refraktor_amd import flame as fl
from refraktor_amd import render

pytestmark = pytest.mark.gpu

W, H = 320, 180


K_F32 = 8   # float32 ulps, as stated for every variation (tests/test_gpu_variations.py)


def _layout(f):
    return oracle.Layout(f, oracle.load_variation_params(render.CONFIG_DIR))


_DB = {}


def _bound(f, st, x0, tr, xid, pass0, precision, K, per_walker_xid=False):
    if "db" not in _DB:
        _DB["db"] = xb.load_db(render.CONFIG_DIR)
    mdb, cm = _DB["db"]
    inf = st.info()
    r = xb.trace_bound(mdb, cm, _layout(f), st.params(), inf.threads_per_block, inf.seed, 0, pass0, x0, tr, xid,
                       precision, K, per_walker_xid=per_walker_xid)
    print(f"[bound] {r}")
    return r


@pytest.fixture(scope="module")
def small(gpu, db):
    f = fl.flame_small(1)
    k = db.compile(f)
    return f, k


def test_warmup_params_match_oracle(small):
    """per-block parameters after evaluation + precalc + jitter"""
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=11, fuse=0, walkers=2048)
    inf = st.info()
    gp = st.params()
    L, op, opal = oracle.prepare(f, render.CONFIG_DIR, W, H, inf.blocks, 11)
    n = L.size_reals
    np.testing.assert_allclose(gp[:, :n], op[:, :n], rtol=2e-6, atol=2e-5)
    assert np.array_equal(st.palettes(), opal)


def test_seed_positions_match_oracle(small):
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=5, fuse=0, walkers=2048)
    inf = st.info()
    g = st.walkers()
    x, y, c = oracle.seed_walkers(st.params(), inf.param_stride, inf.blocks, inf.threads_per_block, W, H, 5)
    np.testing.assert_allclose(g[:, 0], x, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g[:, 1], y, rtol=1e-5, atol=1e-5)
    assert np.array_equal(g[:, 2], c.astype(np.float32))


@pytest.mark.parametrize("maker", [fl.flame_small, fl.flame_1080])
def test_trajectory_one_step(db, gpu, maker):
    """every pass of every walker, fed the GPU's previous state: xform ids
    equal the oracle's, and every step not on a knife edge is within
    K = 8 float32 ulps of the float64 step plus the input-rounding and Monte
    Carlo arithmetic spreads (tests/xform_bound.py)"""
    f = maker(3)
    k = db.compile(f)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=9, fuse=3, walkers=4096)
    inf = st.info()
    x0 = st.walkers()
    P = st.params()
    tr, xid = st.trace(16)
    otr, oxid = oracle.trace(_layout(f), P, inf.param_stride, inf.blocks, inf.threads_per_block, 9, 3, 16,
                             x0[:, 0], x0[:, 1], x0[:, 2], gpu_trace=tr)
    assert np.array_equal(xid, oxid)
    assert np.array_equal(tr[..., 3], otr[..., 3].astype(np.float32))
    ok = tr[..., 3] != 0
    assert np.all(np.abs(tr[..., 2][ok] - otr[..., 2][ok]) <= 1e-6)   # colour blend
    r = _bound(f, st, x0, tr, xid, 3, 0, K_F32)
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["knife_steps"] <= 0.01 * r["steps"], r


def test_trajectory_free_running(small):
    """the first passes from a common start, without feeding GPU state back"""
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=21, fuse=0, walkers=4096)
    inf = st.info()
    x0 = st.walkers()
    tr, xid = st.trace(8)
    otr, oxid = oracle.trace(_layout(f), st.params(), inf.param_stride, inf.blocks, inf.threads_per_block, 21, 0,
                             8, x0[:, 0], x0[:, 1], x0[:, 2])
    assert np.array_equal(xid, oxid)
    err = np.abs(tr[..., :2] - otr[..., :2]) / np.maximum(1.0, np.abs(otr[..., :2]))
    assert np.mean(err[:4] < 1e-3) > 0.99


def _bin_compare(st, pts, label):
    st.clear_histogram()
    st.bin_points(pts)
    g = st.histogram()
    o = oracle.bin_points(st.params()[0, :6], st.palettes()[0], pts, W, H)
    if not np.array_equal(g, o):
        bad = np.argwhere((g != o).any(axis=-1))
        S = st.params()[0, :6].astype(np.float32)
        sx = np.trunc(S[0] * pts[:, 0] + S[2] * pts[:, 1] + S[4])
        sy = np.trunc(S[1] * pts[:, 0] + S[3] * pts[:, 1] + S[5])
        lines = []
        for yy, xx in bad[:5]:
            near = np.nonzero((np.abs(sx - xx) <= 1) & (np.abs(sy - yy) <= 1))[0][:4]
            lines.append(f"pixel ({yy},{xx}) gpu {g[yy, xx]} oracle {o[yy, xx]} points {pts[near].tolist()}")
        pytest.fail(f"{label}: {len(bad)} pixels differ, gpu sum {g.sum(axis=(0, 1))} oracle sum "
                    f"{o.sum(axis=(0, 1))}\n" + "\n".join(lines))


@pytest.mark.parametrize("queue", [False, True])
def test_iterate_draws_equal_bin_points(small, db, queue):
    """the iterate kernel's own draws (palette lerp from its LDS table) add
    exactly what the oracle's bin_points adds for the same points: a traced
    run and a binned run of the same walkers from the same warmup see the
    same points (the small flame has no final xform), binned per block with
    the block's screen transform and palette"""
    f, k = small
    if queue:
        k = db.compile(f, bin_queue=True)
    s = k.pack_samples(W, H)
    passes, walkers = 8, 4096
    st = k.warmup(s, W, H, seed=4, fuse=2, walkers=walkers)
    tr, _ = st.trace(passes)
    st2 = k.warmup(s, W, H, seed=4, fuse=2, walkers=walkers)
    r = st2.bin(iters=walkers * passes)
    assert r["total_passes"] == walkers * passes
    got = st2.histogram()
    inf = st.info()
    P, pal, tpb = st.params(), st.palettes(), inf.threads_per_block
    want = np.zeros_like(got)
    for b in range(inf.blocks):
        want += oracle.bin_points(P[b, :6], pal[b], tr[:, b * tpb:(b + 1) * tpb].reshape(-1, 4), W, H)
    assert int(want[..., 3].sum()) > 0
    np.testing.assert_array_equal(got, want)


def test_bin_points_bit_exact(small):
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=2, fuse=0, walkers=1024)
    rng = np.random.default_rng(0)
    n = 200_000
    inside = np.stack([rng.uniform(-2.0, 2.0, n), rng.uniform(-1.1, 1.1, n), rng.uniform(0.0, 1.0, n),
                       rng.uniform(0.0, 1.0, n)], axis=1).astype(np.float32)
    _bin_compare(st, inside, "in-frame points")
    wide = np.stack([rng.uniform(-2.5, 2.5, n), rng.uniform(-1.5, 1.5, n), rng.uniform(-0.2, 1.2, n),
                     rng.uniform(-0.1, 1.0, n)], axis=1).astype(np.float32)
    _bin_compare(st, wide, "points around the frame, colours and opacities out of range")
    # edge cases: non-finite, zero opacity, exact colour ends
    edge = inside[:60].copy()
    edge[:10, 0] = np.nan
    edge[10:20, 1] = np.inf
    edge[20:30, 3] = 0.0
    edge[30:40, 2] = 1.0
    edge[40:50, 2] = 0.0
    edge[50:60, 2] = np.nan
    _bin_compare(st, edge, "edge cases")


def test_rewarm_matches_fresh_warmup(small):
    """a state re-warmed for a new frame renders what a new state renders"""
    f, k = small
    s = k.pack_samples(W, H)
    fresh = k.warmup(s, W, H, seed=11, fuse=10, walkers=8192)
    fresh.bin(iters=8192 * 32)
    st = k.warmup(s, W, H, seed=3, fuse=7, walkers=8192)
    st.bin(iters=8192 * 5)
    st.rewarm(s, seed=11, fuse=10)
    r = st.bin(iters=8192 * 32)
    assert r["total_passes"] == 8192 * 32
    np.testing.assert_array_equal(st.histogram(), fresh.histogram())
    np.testing.assert_array_equal(st.walkers(), fresh.walkers())


def test_render_counters_and_determinism(small):
    f, k = small
    out = []
    for _ in range(2):
        st = k.warmup(k.pack_samples(W, H), W, H, seed=4, fuse=10, walkers=8192)
        res = st.bin(iters=8192 * 64)
        h = st.histogram()
        assert res["total_passes"] == 8192 * 64
        assert int(h[..., 3].astype(np.uint64).sum()) == res["hit_sum"]
        out.append(h)
    assert np.array_equal(out[0], out[1]), "integer binning with counter RNG must be reproducible"


@pytest.mark.parametrize("flame_seed", [3, 6])
def test_c1_converged_render_matches_oracle(db, gpu, flame_seed):
    """The C1 workload (640x360 flame of linear / sinusoidal / spherical
    xforms, 2^16 walkers x 1024 passes + 20 fuse = 2^26 iterations) against
    the oracle's host render at the same budget: the per-pixel relative L1
    over pixels with >= 256 draws, from independent streams, within TAU of
    the oracle's own seed-to-seed figure (tests/parity_stats.py)."""
    w, h, walkers, passes = 640, 360, 1 << 16, 1024
    f = fl.flame_small(flame_seed)
    k = db.compile(f)
    st = k.warmup(k.pack_samples(w, h), w, h, seed=13, fuse=20, walkers=walkers)
    res = st.bin(iters=walkers * passes)
    assert res["total_passes"] == walkers * passes
    g = st.histogram()
    assert int(g[..., 3].astype(np.uint64).sum()) == res["hit_sum"]
    inf = st.info()
    L = _layout(f)
    ref, _ = oracle.render(L, st.params(), inf.param_stride, st.palettes(), inf.blocks, inf.threads_per_block, w, h,
                           14, 20, passes)
    other, _ = oracle.render(L, st.params(), inf.param_stride, st.palettes(), inf.blocks, inf.threads_per_block, w,
                             h, 15, 20, passes)
    ratio, s_gpu, s_seed, n = converged_ratio(g, ref, other)
    print(f"[c1 seed {flame_seed}] pixels {n} gpu {s_gpu:.4f} oracle seeds {s_seed:.4f} ratio {ratio:.3f}")
    assert n > 10000
    assert ratio <= TAU, (ratio, s_gpu, s_seed)


def test_tonemap_matches_oracle(small):
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=3, fuse=10, walkers=8192)
    st.bin(iters=8192 * 64)
    h = st.histogram()
    q = st.info().quality
    for fmt in (0, 1, 2, 3):
        img = st.tonemap(2.2, 4.0, 1.0, fmt=fmt).astype(np.float64)
        ref = oracle.tonemap(h, q, 2.2, 4.0, 1.0)
        ch = img.shape[-1]
        tol = 2e-3 if fmt in (1, 3) else 2e-5
        np.testing.assert_allclose(img, ref[..., :ch], atol=tol)


def test_density_histogram_matches_oracle(small):
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=3, fuse=10, walkers=8192)
    st.bin(iters=8192 * 64)
    bins, mx = st.density_histogram()
    ob, omx = oracle.density_histogram(st.histogram())
    assert np.array_equal(bins, ob) and mx == omx


def test_xaos_trajectory_one_step(db, gpu):
    """per-walker xaos picks (waterfall over the wave's xforms) vs the oracle"""
    f = fl.flame_xaos(2)
    k = db.compile(f)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=9, fuse=3, walkers=4096)
    inf = st.info()
    x0, last0 = st.walkers(), st.walkers_last()
    assert (last0 < 5).all()  # the fuse passes picked an xform for everyone
    tr, xid = st.trace(16)
    assert xid.shape == (16, inf.walkers)
    otr, oxid = oracle.trace(_layout(f), st.params(), inf.param_stride, inf.blocks, inf.threads_per_block, 9, 3, 16,
                             x0[:, 0], x0[:, 1], x0[:, 2], gpu_trace=tr, last=last0, gpu_xids=xid)
    assert np.array_equal(xid, oxid)
    # rows with zero entries are never taken: xform j after i only if chaos[i][j] != 0
    ch = np.array(f["chaos"])
    prev = np.concatenate([last0[None, :], xid[:-1]])
    assert (ch[prev, xid] != 0).all()
    r = _bound(f, st, x0, tr, xid, 3, 0, K_F32, per_walker_xid=True)
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["knife_steps"] <= 0.01 * r["steps"], r


def _select_f64(P, L, ratio):
    """the f64 kernel's xform pick: cumulative weights in double
    (flame.tpl:110-123) with rand01 = u / 4294967295 in double"""
    ratio = ratio * P[12]
    rsum, last_nonzero = 0.0, 0
    n = L.n_xforms
    for i in range(n):
        w = P[L.xforms[i]["weight"]]
        if i + 1 < n:
            if w != 0.0 and rsum + w >= ratio:
                return i
            rsum += w
            if w != 0.0:
                last_nonzero = i
        else:
            return i if w != 0.0 else last_nonzero
    return last_nonzero


def test_f64_trajectory_one_step(db, gpu):
    """double-precision kernel from its own double parameters: xform ids from
    the double pick, every step within K = 1 float32 ulp (the walker store
    rounds to float32) plus the double-precision spreads"""
    f = fl.flame_1080(5)
    k = db.compile(f, precision=1)
    st = k.warmup(k.pack_samples(W, H), W, H, seed=4, fuse=3, walkers=2048)
    inf = st.info()
    x0 = st.walkers()
    tr, xid = st.trace(8)
    P = st.params()
    assert P.dtype == np.float64
    L = _layout(f)
    waves = inf.threads_per_block // 64
    for i in range(8):
        pass_ = 3 + i
        ctr = np.array([[pass_ >> 2, b, w, oracle.TAG_PICK] for b in range(inf.blocks) for w in range(waves)],
                       dtype=np.uint32)
        word = oracle.philox(ctr, 4, 0)[:, pass_ & 3]
        want = [_select_f64(P[j // waves], L, float(u) / 4294967295.0) for j, u in enumerate(word)]
        assert np.array_equal(xid[i], np.array(want, dtype=np.uint32)), i
    r = _bound(f, st, x0, tr, xid, 3, 1, 1)
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["knife_steps"] <= 0.01 * r["steps"], r


def test_animated_warmup_params_match_oracle(db, gpu):
    """Catmull-Rom over four different keyframe samples, per-block times and polar affines"""
    kf = [fl.flame_keyframe(40 + i) for i in range(2)]
    for k_ in kf[1:]:
        k_["palette"] = kf[0]["palette"]
    aligned = db.align(kf)
    k = db.compile(aligned[0])
    packed = [a.pack_samples(W, H, 0.0, 0.0, 1) for a in aligned]
    samples = np.concatenate([packed[0], packed[0], packed[1], packed[1]])
    st = k.warmup(samples, W, H, seed=6, fuse=0, walkers=64 * 256)
    inf = st.info()
    L = oracle.Layout(aligned[0].to_json(), oracle.load_variation_params(render.CONFIG_DIR))
    ss = L.size_reals + 768
    fixed = oracle.fix_rotation(samples.copy(), ss, L.size_reals, L.affine_indices, 4)
    seg = oracle.catmull(fixed, ss)
    P, pals = oracle.eval_samples(seg, L, inf.blocks)
    P = oracle.precalc(P, L, 6, 0)
    n = L.size_reals
    np.testing.assert_allclose(st.params()[:, :n], P[:, :n], rtol=2e-6, atol=2e-5)
    assert np.array_equal(st.palettes(), pals)
    # the weights really move between the two keyframes
    w0 = L.xforms[0]["weight"]
    assert abs(float(P[0, w0]) - float(P[-1, w0])) > 1e-3


def test_density_estimate_matches_oracle(small):
    f, k = small
    st = k.warmup(k.pack_samples(96, 64), 96, 64, seed=3, fuse=10, walkers=4096)
    st.bin(iters=4096 * 16)
    h = st.histogram()
    # radii of the 1080p (0-9) and 4K (0-12) configurations included
    for max_r, min_r, curve in ((0.0, 0.0, 0.4), (3.0, 0.0, 0.4), (6.0, 1.0, 0.6), (9.0, 0.0, 0.4), (12.0, 0.5, 0.4)):
        got = st.density_estimate(max_r, min_r, curve).astype(np.float64)
        want = oracle.density_estimate(h, max_r, min_r, curve)
        np.testing.assert_allclose(got, want, rtol=2e-4, atol=1e-4 * want.max())


def test_bailouts(small):
    f, k = small
    s = k.pack_samples(W, H)
    # exact iteration budget, rounded up to whole passes of all walkers
    st = k.warmup(s, W, H, seed=1, fuse=5, walkers=8192)
    r = st.bin(iters=8192 * 10 + 1)
    assert r["total_passes"] == 8192 * 11
    # quality target: every block stops at its share of the target density
    st = k.warmup(s, W, H, seed=1, fuse=5, walkers=8192)
    r = st.bin(quality=4.0)
    assert r["quality"] >= 4.0 * 0.99 and r["quality"] < 8.0
    # time limit on the device clock
    st = k.warmup(s, W, H, seed=1, fuse=5, walkers=8192)
    r = st.bin(millis=3.0)
    assert r["total_passes"] > 0 and r["elapsed_ms"] < 100.0
    # abort from the host while the kernel runs (no other bailout: it would run ~forever)
    st = k.warmup(s, W, H, seed=1, fuse=5, walkers=8192)
    st.bin_async(millis=60000.0)
    import time
    time.sleep(0.05)
    st.abort()
    r = st.bin_wait()
    assert r["total_passes"] > 0 and r["kernel_ms"] < 30000.0


def test_temporal_multiplier_blocks(db, gpu):
    """tm blocks per base block, each its own temporal sample along the segment"""
    kf = [fl.flame_keyframe(60), fl.flame_keyframe(61)]
    kf[1]["palette"] = kf[0]["palette"]
    aligned = db.align(kf)
    k = db.compile(aligned[0])
    packed = [a.pack_samples(W, H, 0.0, 0.0, 1) for a in aligned]
    samples = np.concatenate([packed[0], packed[0], packed[1], packed[1]])
    st = k.warmup(samples, W, H, seed=2, fuse=5, walkers=4096, temporal_multiplier=4)
    inf = st.info()
    assert inf.blocks == 4 * 4096 // inf.threads_per_block
    P = st.params()
    assert len({tuple(np.round(r[:12], 5)) for r in P}) > 1  # rows differ along the segment
    r = st.bin(iters=inf.walkers * 8)
    assert r["total_passes"] == inf.walkers * 8


def test_density_hdr_matches_oracle(small):
    f, k = small
    st = k.warmup(k.pack_samples(W, H), W, H, seed=3, fuse=10, walkers=8192)
    st.bin(iters=8192 * 64)
    img = st.tonemap(2.2, 4.0, 1.0)
    bins, _ = st.density_histogram()
    got = render.density_hdr(img, bins)
    want = oracle.density_hdr(img, bins)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)


def test_density_estimate_paths_agree(small):
    """a tile with fewer non-empty sources than window offsets visits its
    source list, a crowded tile its window: both give the same bits. Dense
    sources (kernel radius below one pixel: they reach only their own pixel)
    fill the free pixels of the left half in B; every other pixel must come
    out as in A"""
    import ctypes as C
    from refraktor_amd import _lib
    f, k = small
    w, h = 64, 48
    st = k.warmup(k.pack_samples(w, h), w, h, seed=3, fuse=4, walkers=1024)
    inf = st.info()
    assert inf.hist_bits == 32
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:h, 0:w]
    sparse = rng.random((h, w)) < 0.05
    A = np.zeros((h, w, 4), np.uint32)
    A[sparse, :3] = rng.integers(0, 256, size=(int(sparse.sum()), 3))
    A[sparse, 3] = 255  # one full-opacity draw: the widest kernel
    dense = ~sparse & (xx < 32)
    B = A.copy()
    B[dense, :3] = rng.integers(0, 256, size=(int(dense.sum()), 3)) * 4096
    B[dense, 3] = 255 * 4096  # 4096 draws: kernel radius 9 / 4096^0.4 < 1
    outs = []
    for hist in (A, B):
        hist = np.ascontiguousarray(hist)
        _lib.check(_lib.lib().rfk_memcpy_htod(C.c_void_p(inf.histogram), hist.ctypes.data_as(C.c_void_p), hist.nbytes))
        outs.append(st.density_estimate(9.0))
    # the span of the first 16x16 tile (radius 9 around it) holds fewer
    # sources than the 19x19 window in A and more in B
    span_a = int(sparse[:25, :25].sum())
    span_b = span_a + int(dense[:25, :25].sum())
    assert span_a < 19 * 19 < span_b, (span_a, span_b)
    keep = ~dense
    np.testing.assert_array_equal(outs[0][keep], outs[1][keep])
    assert (outs[0][keep] != 0).any()


def test_animated_flame_json_matches_oracle(db, gpu):
    """a flame JSON with increase / sine / mix animations (rotation, scale,
    xform weight and colour, vlink rotation, affine term, variation weight and
    parameter), packed at four times by the library: the per-block parameters
    after warmup equal the oracle's packing, Catmull-Rom and precalc of the
    same times, and one-step trajectories with those per-block parameters
    keep the stated bound"""
    from test_cpu_anima import animated_flame
    f = animated_flame()
    fo = db.flame(f)
    k = db.compile(fo)
    samples = fo.pack_samples(W, H, 0.0, 0.25, 4)
    st = k.warmup(samples, W, H, seed=6, fuse=3, walkers=64 * 256)
    inf = st.info()
    js = fo.to_json()
    L = oracle.Layout(js, oracle.load_variation_params(render.CONFIG_DIR))
    osamp = oracle.pack_samples(f, L, W, H, 0.0, 0.25, 4)
    np.testing.assert_allclose(samples, osamp, rtol=1e-12, atol=1e-12)
    ss = L.size_reals + 768
    fixed = oracle.fix_rotation(osamp.copy(), ss, L.size_reals, L.affine_indices, 4)
    P, pals = oracle.eval_samples(oracle.catmull(fixed, ss), L, inf.blocks)
    P = oracle.precalc(P, L, 6, 0)
    n = L.size_reals
    np.testing.assert_allclose(st.params()[:, :n], P[:, :n], rtol=2e-6, atol=2e-5)
    assert np.array_equal(st.palettes(), pals)
    assert not np.allclose(st.params()[0, 12:n], st.params()[-1, 12:n])   # blocks see different times
    x0 = st.walkers()
    tr, xid = st.trace(6)
    r = _bound(js, st, x0, tr, xid, 3, 0, K_F32)
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["knife_steps"] <= 0.01 * r["steps"], r

"""Batched animation: one persistent launch (optionally a hipGraph replay)
against per-frame renders through warmup/bin with the same seeds.

Frame f of a batch whose frame count equals its segment count renders the
Catmull-Rom segment f with block b at t = b / blocks, which is exactly what a
single render of control samples f..f+3 does; integer binning makes the
histograms comparable bit for bit.
"""
import numpy as np
import pytest

from refraktor_amd import animation
from refraktor_amd import flame as fl
from refraktor_amd import _lib
from refraktor_amd import render

pytestmark = pytest.mark.gpu

W, H = 320, 180
WALKERS = 1 << 14


def _keyframes(n: int, seed: int) -> list[dict]:
    kf = animation.keyframe_flames(n, seed)
    # one palette for all keyframes keeps the hue curve free of 360-degree
    # unwrapping, so a 4-sample slice of the control samples is unwrapped
    # exactly as the whole sequence is
    for k in kf[1:]:
        k["palette"] = kf[0]["palette"]
    return kf


def test_batch_frames_match_single_renders(gpu, db):
    kf = _keyframes(5, 3)
    aligned, samples = animation.control_samples(db, kf, W, H)
    k = db.compile(aligned[0])
    nseg = len(kf) - 1
    ss = samples.size // (len(kf) + 2)
    iters = np.array([1 << 20, 3 << 19, 1 << 19, 5 << 18], dtype=np.uint64)[:nseg]
    b = k.batch(samples, nseg, iters, W, H, walkers=WALKERS, seed=9, fuse=12)
    assert b.items == nseg * (WALKERS // k.threads_per_block)
    b.run(graph=False)
    for f in range(nseg):
        info = b.frame(f)
        st = k.warmup(samples[f * ss:(f + 4) * ss], W, H, seed=9, fuse=12, walkers=WALKERS, stream_id=f)
        res = st.bin(iters=int(iters[f]))
        assert info["total_passes"] == res["total_passes"]
        assert info["total_draws"] == res["total_draws"]
        assert info["bad_values"] == res["bad_values"]
        np.testing.assert_array_equal(b.histogram(f), st.histogram())


def test_batch_graph_replays_match_eager(gpu, db):
    kf = _keyframes(4, 5)
    anim = animation.Animation(db, kf, W, H, frames=9, iters=animation.frame_budgets(9, 4, 1 << 18, 1 << 21),
                               walkers=WALKERS, seed=2)
    anim.render(graph=False)
    eager = [anim.batch.histogram(f) for f in range(9)]
    # frames differ (own streams, own times) and each got its budget
    assert not np.array_equal(eager[0], eager[1])
    for f in range(9):
        inf = anim.batch.frame(f)
        assert inf["total_passes"] == inf["passes_per_walker"] * inf["walkers"]
        assert inf["passes_per_walker"] * inf["walkers"] >= int(anim.iters[f])
        assert 0 < inf["total_draws"] <= inf["total_passes"]
        assert int(eager[f][..., 3].sum(dtype=np.uint64)) > 0
    for rep in range(3):
        anim.render(graph=True)
        for f in range(9):
            np.testing.assert_array_equal(anim.batch.histogram(f), eager[f], err_msg=f"replay {rep} frame {f}")
        # unrelated eager work between replays
        st = db.compile(fl.flame_small(1)).warmup(db.flame(fl.flame_small(1)).pack_samples(64, 64), 64, 64, seed=1,
                                                   walkers=1 << 12)
        st.bin(iters=1 << 16)


def test_queue_batch_matches_atomic_batch(gpu, db):
    """queue binning cuts each frame into slices of passes and runs several
    frames' slices per chunk launch; integer sums make every frame equal to
    the atomic batch's, bit for bit, eagerly and in graph replays"""
    kf = _keyframes(4, 6)
    n, walkers = 12, 1 << 16
    iters = animation.frame_budgets(n, 5, 1 << 23, 1 << 26)
    atomic = animation.Animation(db, kf, W, H, frames=n, iters=iters, walkers=walkers, seed=3)
    queued = animation.Animation(db, kf, W, H, frames=n, iters=iters, walkers=walkers, seed=3, bin_queue=True)
    bpf = walkers // queued.kernel.threads_per_block
    # more frames than slots per chunk, and frames of several slices
    assert queued.batch.grid < n * bpf
    assert queued.batch.items > n * bpf
    atomic.render(graph=False)
    queued.render(graph=False)
    eager = []
    for f in range(n):
        a, q = atomic.batch.frame(f), queued.batch.frame(f)
        for key in ("total_passes", "total_draws", "bad_values", "passes_per_walker", "walkers"):
            assert a[key] == q[key], (f, key)
        eager.append(queued.batch.histogram(f))
        np.testing.assert_array_equal(eager[f], atomic.batch.histogram(f), err_msg=f"frame {f}")
    for rep in range(2):
        queued.render(graph=True)
        for f in range(n):
            np.testing.assert_array_equal(queued.batch.histogram(f), eager[f], err_msg=f"replay {rep} frame {f}")


def test_queue_batch_64bit_bins(gpu, db):
    """the queue batch's per-frame histogram offsets in 64-bit bins (the
    converged-render format): every frame equal to the atomic 64-bit batch"""
    kf = _keyframes(4, 8)
    aligned, samples = animation.control_samples(db, kf, W, H)
    n, walkers = 10, 1 << 16
    iters = animation.frame_budgets(n, 9, 1 << 23, 1 << 25)
    ka = db.compile(aligned[0], hist_bits=64)
    kq = db.compile(aligned[0], hist_bits=64, bin_queue=True)
    a = ka.batch(samples, n, iters, W, H, walkers=walkers, seed=4)
    q = kq.batch(samples, n, iters, W, H, walkers=walkers, seed=4)
    a.run(graph=False)
    q.run(graph=True)
    for f in range(n):
        assert q.frame(f)["hist_bits"] == 64
        hq = q.histogram(f)
        assert hq.dtype == np.uint64 and int(hq[..., 3].sum()) > 0
        np.testing.assert_array_equal(hq, a.histogram(f), err_msg=f"frame {f}")


def test_batch_rejects_bad_arguments(gpu, db):
    k = db.compile(fl.flame_keyframe(1))
    s = k.pack_samples(W, H)
    with pytest.raises(_lib.RfkError):
        k.batch(s[: s.size // 2], 2, 1 << 20, W, H)  # fewer than 4 samples
    with pytest.raises(_lib.RfkError):
        k.batch(s, 2, 0, W, H)  # empty budget

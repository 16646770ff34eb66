"""Every variation of the database on the GPU, one step, against the oracle's
interpreter of the same program (oracle/flang_eval.py, float64), which the CPU
tests pin to librefrakt's 99 programs through the golden vectors.

One flame per parameter set of the fixture: 99 xforms, each one vlink with
the set's affine and one variation at the set's weight and parameters. The
walkers are set to chosen points, one pass is traced, and each walker's
result is compared with the interpreter evaluated at the walker's point
after the vlink affine, with the variation parameters the GPU itself holds
(its precalc checked separately) and the walker's own Philox lane words
(counter {pass, walker, k, LANE}, key {seed, stream}), so random variations
are compared draw for draw.

The bound, per output coordinate:
    |gpu - ref| <= K * 2^-23 * max(1, |ref|) + 2 * s + 2 * n
where s is the spread of the float64 program over the rounding of its input
(the point moved by 4 ulps of the affine's terms along each axis, float32
ulps for the f32 kernel, double ulps for the f64 one), and n the spread of
MCA_RUNS Monte Carlo arithmetic runs of the program (oracle/flang_eval.py
Perturb: every operation perturbed by 2^-24 relative, every library result
by 2^-21 relative and, for the periodic functions, 2^-21 absolute, which is
the accuracy of the f32 kernels' hardware transcendentals; the f64 kernel's
runs use its double-precision figures). s and n are what the arithmetic
itself allows near poles, cancellations and steep slopes; K is the stated
allowance on top, in float32 ulps of max(1, |ref|): 8 for the f32 kernel
(K_F32) and 1 for the f64 one, whose float32 walker store rounds the result
(K_F64). The check
keeps its power: the median bound over the generic points (the uniform ones,
not those placed on the axes or near the origin) stays within POWER_MAX of
the result's scale. Knife edges, where the program jumps by more than 1e-3 of its scale
within s or n (branch cuts, floor, comparisons), stay under KNIFE_EDGE_MAX
of the generic points and are not compared. The GPU resets a walker
(opacity 0) only where a perturbed run gives a bad value (|v| > 1e9 or NaN)
and always where every run and the exact result do.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import flang_eval as fe
from oracle import oracle
from refraktor_amd import _lib
from refraktor_amd import flame as fl
from refraktor_amd import render

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
FIXTURE = ROOT / "tests" / "golden" / "variation_vectors.npz"
ULP = 2.0 ** -23
WALKERS = 1 << 17
SEED = 23

# K in float32 ulps of max(1, |result|): one figure for all 99 variations
# (the largest any needed when it was set was 3.2, gaussian_blur in f32)
K_F32 = 8
K_F64 = 1            # f64 kernel: its float32 walker store rounds the result
KNIFE_EDGE_MAX = 0.05
POWER_MAX = 1e-4     # median bound over generic points, relative to max(1, |result|)
MCA_RUNS = 8


@pytest.fixture(scope="module")
def vec():
    z = np.load(FIXTURE)
    return {"names": [str(n) for n in z["names"]], "params": json.loads(str(z["params"])),
            "weights": z["weights"], "affines": z["affines"]}


@pytest.fixture(scope="module")
def mine():
    cfg = Path(render.CONFIG_DIR)
    return fe.load_db((cfg / "variations_new.yml").read_text(), (cfg / "common.yml").read_text())


def _flame(vec, s):
    a, b, c, d, e, f = vec["affines"][s]
    n = len(vec["names"])
    xfs = []
    for i, name in enumerate(vec["names"]):
        prm = vec["params"][name][s]
        w = float(vec["weights"][s])
        ve = {"weight": w, "parameters": prm} if prm else w
        xfs.append(fl.xform(1.0, i / (n - 1), 0.5, [fl.vlink([a, d, b, e, c, f], {name: ve})]))
    return fl.make_flame(xfs, [[0.0, 0.0, 1.0]] * 256)


def _inputs(rng, aff, n):
    """pre-images under the affine of points spread over [-2.5, 2.5]^2, with a
    share near the origin and on the axes (after the affine)"""
    a, b, c, d, e, f = aff
    px = rng.uniform(-2.5, 2.5, n)
    py = rng.uniform(-2.5, 2.5, n)
    k = n // 32
    px[:k] *= 1e-3
    py[:k] *= 1e-3
    px[k:2 * k] = 0.0
    py[2 * k:3 * k] = 0.0
    A = np.array([[a, b], [d, e]])
    pre = np.linalg.solve(A, np.stack([px - c, py - f]))
    return pre[0].astype(np.float32), pre[1].astype(np.float32)


def _lane_words(gids, seed, stream=0, pass_=0, nwords=16):
    out = np.zeros((gids.size, nwords), dtype=np.uint64)
    for k in range(nwords // 4):
        ctr = np.stack([np.full_like(gids, pass_), gids, np.full_like(gids, k),
                        np.full_like(gids, oracle.TAG_LANE)], axis=1).astype(np.uint32)
        out[:, 4 * k:4 * k + 4] = oracle.philox(ctr, seed, stream)
    return out


def _affine(P, off, x, y):
    a, d, b, e, c, f = (float(v) for v in P[off:off + 6])
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    px, py = a * xd + b * yd + c, d * xd + e * yd + f
    scale_x = np.abs(a * xd) + np.abs(b * yd) + abs(c)
    scale_y = np.abs(d * xd) + np.abs(e * yd) + abs(f)
    return px, py, {"a": a, "b": b, "c": c, "d": d, "e": e, "f": f}, scale_x, scale_y


def _jump(q, r):
    """|q - r|, 0 where both are the same non-finite value, inf where only
    one of them is NaN"""
    with np.errstate(invalid="ignore"):
        d = np.where(q == r, 0.0, np.abs(q - r))
        both_nan = np.isnan(q) & np.isnan(r)
        return np.where(both_nan, 0.0, np.where(np.isnan(d), np.inf, d))


def _run(db, vec, mine, s, precision):
    mdb, commons = mine
    fd = _flame(vec, s)
    k = db.compile(fd, precision=precision)
    st = k.warmup(k.pack_samples(64, 64), 64, 64, seed=SEED, fuse=0, walkers=WALKERS)
    inf = st.info()
    rng = np.random.default_rng(100 + s)
    x, y = _inputs(rng, vec["affines"][s], inf.walkers)
    c = rng.uniform(0.0, 1.0, inf.walkers).astype(np.float32)
    for ptr, arr in ((inf.walker_x, x), (inf.walker_y, y), (inf.walker_c, c)):
        _lib.check(_lib.lib().rfk_memcpy_htod(ptr, arr.ctypes.data, arr.nbytes))
    tr, xid = st.trace(1)
    P = st.params()[0].astype(np.float64)
    L = oracle.Layout(fd, oracle.load_variation_params(render.CONFIG_DIR))
    wave_of = np.arange(inf.walkers) // 64
    xw = xid[0][wave_of]
    words = _lane_words(np.arange(inf.walkers, dtype=np.uint64), SEED)
    report = {}
    for i, name in enumerate(vec["names"]):
        m = np.nonzero(xw == i)[0]
        if m.size == 0:
            report[name] = None
            continue
        xs = L.xforms[i]
        vs = xs["vlinks"][0]
        slot = vs["vars"][0]
        px, py, aff, sx, sy = _affine(P, vs["aff"], x[m], y[m])
        w = float(P[slot["weight"]])
        prm = {pn: float(P[off]) for pn, off in slot["params"].items()}
        v = mdb[name]
        # the GPU's precalc against the interpreter's, from the same inputs
        user = {pn: prm[pn] for pn in v.user_params}
        full = fe.precalc(v, user, aff)
        for pn, off in slot["params"].items():
            if v.params[pn]["precalc"] and np.isfinite(full[pn]):
                tol = (1e-5 if precision == 0 else 1e-12) * max(1.0, abs(full[pn]))
                assert abs(prm[pn] - full[pn]) <= tol, (name, pn, prm[pn], full[pn])
        wd = words[m]

        def ev(qx_, qy_, mca=None):
            rng_ = fe.Random(wd, double_divide=precision == 1)
            return fe.evaluate(v, commons, qx_, qy_, w, prm, aff, rng=rng_, with_precalc=False, mca=mca)

        rx, ry = ev(px, py)
        spread_x = np.zeros_like(rx)
        spread_y = np.zeros_like(ry)
        # the input's rounding: the f32 kernel forms p in float32, the f64
        # kernel in double from float32 walkers
        ulp_in = ULP if precision == 0 else 2.0 ** -52
        dx, dy = 4 * ulp_in * np.maximum(sx, 1e-30), 4 * ulp_in * np.maximum(sy, 1e-30)
        for ox, oy in ((dx, 0), (-dx, 0), (0, dy), (0, -dy)):
            qx, qy = ev(px + ox, py + oy)
            spread_x = np.maximum(spread_x, _jump(qx, rx))
            spread_y = np.maximum(spread_y, _jump(qy, ry))
        # Monte Carlo arithmetic: the program with every operation perturbed at
        # the kernel's precision and every library result at its functions'
        # accuracy (f32: hardware transcendentals; f64: the library's)
        mrng = np.random.default_rng(1000 + 7 * i + s)
        if precision == 0:
            pert = dict(op_rel=2.0 ** -24, fn_rel=2.0 ** -21, fn_abs=2.0 ** -21)
        else:
            pert = dict(op_rel=2.0 ** -53, fn_rel=2.0 ** -51, fn_abs=2.0 ** -51)
        nx, ny = np.zeros_like(rx), np.zeros_like(ry)
        any_bad = np.zeros(rx.shape, dtype=bool)
        any_ok = np.zeros(rx.shape, dtype=bool)
        for _ in range(MCA_RUNS):
            qx, qy = ev(px, py, fe.Perturb(mrng, **pert))
            nx = np.maximum(nx, _jump(qx, rx))
            ny = np.maximum(ny, _jump(qy, ry))
            with np.errstate(invalid="ignore"):
                okq = (np.abs(qx) <= 1e9) & (np.abs(qy) <= 1e9)
            any_ok |= okq
            any_bad |= ~okq
        g = tr[0][m].astype(np.float64)
        generic = m >= 3 * (inf.walkers // 32)   # not one of the placed hard points
        report[name] = (g, rx, ry, spread_x, spread_y, nx, ny, any_bad, any_ok, c[m], xs, P, generic)
    return report, L


def _check(report, vec, K_of, label):
    rows, failures = [], []
    for i, name in enumerate(vec["names"]):
        r = report[name]
        assert r is not None, f"{name}: no wave picked its xform"
        g, rx, ry, sx_, sy_, nx_, ny_, any_bad, any_ok, c_in, xs, P, generic = r
        gx, gy, gc, gop = g[:, 0], g[:, 1], g[:, 2], g[:, 3]
        # colour blend (flamelib.h:21 INTERP) and opacity
        col, speed = float(P[xs["color"]]), float(P[xs["color_speed"]])
        cexp = c_in.astype(np.float64) * (1.0 - speed) + col * speed
        reset = gop == 0.0
        assert np.all(np.abs(gc[~reset] - cexp[~reset]) <= 1e-6), (name, "colour")
        assert np.all(gop[~reset] == 1.0), (name, "opacity")
        K = K_of(name)
        with np.errstate(invalid="ignore", over="ignore"):
            big = np.maximum(1.0, np.maximum(np.abs(rx), np.abs(ry)))
            ref_ok = (np.abs(rx) <= 1e9) & (np.abs(ry) <= 1e9)
            # knife edges: the program jumps within the input's rounding or
            # within the perturbed runs (branch cuts, floor/trunc/rint,
            # comparisons)
            knife = ~(np.maximum(np.maximum(sx_, sy_), np.maximum(nx_, ny_)) <= 1e-3 * big)
            tol_x = K * ULP * np.maximum(1.0, np.abs(rx)) + 2.0 * sx_ + 2.0 * nx_
            tol_y = K * ULP * np.maximum(1.0, np.abs(ry)) + 2.0 * sy_ + 2.0 * ny_
            ex, ey = np.abs(gx - rx), np.abs(gy - ry)
        checked = ~reset & ref_ok & ~knife
        bad = checked & ~((ex <= tol_x) & (ey <= tol_y))
        # resets (opacity 0) are the walkers whose result is a bad value
        # (|v| > 1e10 or NaN): allowed wherever a perturbed run gives one
        wrong_reset = reset & ~any_bad & ~knife
        missed_reset = ~reset & ~ref_ok & ~any_ok & ~knife
        with np.errstate(invalid="ignore"):
            # the K each point needs: its error beyond the arithmetic's share
            need = np.where(checked, np.maximum((ex - 2.0 * sx_ - 2.0 * nx_) / np.maximum(1.0, np.abs(rx)),
                                                (ey - 2.0 * sy_ - 2.0 * ny_) / np.maximum(1.0, np.abs(ry))) / ULP,
                            -np.inf)
            # the check's power: the bound relative to the result's scale
            rel_tol = np.maximum(tol_x / np.maximum(1.0, np.abs(rx)), tol_y / np.maximum(1.0, np.abs(ry)))
        knife_share = float(knife[generic].mean())
        power = float(np.median(rel_tol[checked & generic]))
        rows.append((name, int(g.shape[0]), float(need.max()), knife_share, power, int(bad.sum()),
                     int(wrong_reset.sum()), int(missed_reset.sum())))
        if bad.any() or wrong_reset.any() or missed_reset.any() or knife_share > KNIFE_EDGE_MAX \
                or not power <= POWER_MAX:
            j = int(np.nonzero(bad)[0][0]) if bad.any() else -1
            failures.append((name, rows[-1][2:], None if j < 0 else (float(gx[j]), float(rx[j]), float(gy[j]),
                                                                    float(ry[j]))))
    print(f"\n[{label}] variation points K-needed knife-share median-relative-bound bad wrong-reset "
          "missed-reset")
    for row in rows:
        print(f"[{label}] {row[0]:14s} {row[1]:6d} {row[2]:8.2f} {row[3]:.4f} {row[4]:.3e} {row[5]} {row[6]} "
              f"{row[7]}")
    assert not failures, failures


@pytest.mark.parametrize("s", [0, 1, 2])
def test_every_variation_one_step_f32(db, gpu, vec, mine, s):
    report, _ = _run(db, vec, mine, s, 0)
    _check(report, vec, lambda n: K_F32, f"f32 set {s}")


def test_every_variation_one_step_f64(db, gpu, vec, mine):
    report, _ = _run(db, vec, mine, 0, 1)
    _check(report, vec, lambda n: K_F64, "f64 set 0")

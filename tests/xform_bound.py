"""One chaos-game step of whole xforms, evaluated by the oracle's interpreter
(oracle/flang_eval.py) with the bound the per-variation GPU test states:

    |gpu - ref| <= K * 2^-23 * max(1, |ref|) + 2 * s + 2 * n

ref: the step in float64 from the GPU's input point with the parameters the
GPU holds: each vlink's affine (x' = a x + b y + c, y' = d x + e y + f),
then the sum of its variations in sorted name order (the reference keeps a
vlink's variations in a std::map), every draw taken from the walker's lane
stream in that order; s: the spread over the input moved by 4 ulps of the
affine's terms (float32 ulps where the input is a float32 value the kernel
held in double: the f64 kernel's passes after the first, fed back from its
float32 trace); n: the spread of Monte Carlo arithmetic runs
(flang_eval.Perturb, at the kernel's precision). Knife edges (a jump of more
than 1e-3 of the scale within s or n) are not compared and must stay rare;
a walker the GPU resets (opacity 0) must be one whose step some perturbed
run takes to a bad value.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import flang_eval as fe
from oracle import oracle

ULP = 2.0 ** -23
MCA_RUNS = 6


def load_db(config_dir) -> tuple[dict, dict]:
    cfg = Path(config_dir)
    return fe.load_db((cfg / "variations_new.yml").read_text(), (cfg / "common.yml").read_text())


def jump(q, r):
    """|q - r|, 0 where both are the same non-finite value, inf where only
    one of them is NaN"""
    with np.errstate(invalid="ignore"):
        d = np.where(q == r, 0.0, np.abs(q - r))
        both_nan = np.isnan(q) & np.isnan(r)
        return np.where(both_nan, 0.0, np.where(np.isnan(d), np.inf, d))


def lane_words(gids, seed, stream, pass_, nwords=16):
    """the walkers' lane words of one pass: philox({pass, gid, k, LANE})"""
    gids = np.asarray(gids, dtype=np.uint64)
    out = np.zeros((gids.size, nwords), dtype=np.uint64)
    for k in range(nwords // 4):
        ctr = np.stack([np.full_like(gids, pass_), gids, np.full_like(gids, k),
                        np.full_like(gids, oracle.TAG_LANE)], axis=1).astype(np.uint32)
        out[:, 4 * k:4 * k + 4] = oracle.philox(ctr, seed, stream)
    return out


def _xform_eval(mdb, commons, L, P, xi, x, y, words, precision, offset=(0.0, 0.0), mca=None):
    """(x', y', scale_x, scale_y) of xform xi at float64 points, one row of
    parameters P (float64 copy of the GPU's)"""
    xs = L.xforms[xi]
    rng = fe.Random(words, double_divide=precision == 1)
    ix, iy = x, y
    sx = sy = None
    for vi, vs in enumerate(xs["vlinks"]):
        a, d, b, e, c, f = (float(v) for v in P[vs["aff"]:vs["aff"] + 6])
        aff = {"a": a, "b": b, "c": c, "d": d, "e": e, "f": f}
        px, py = a * ix + b * iy + c, d * ix + e * iy + f
        if vi == 0:
            sx = np.abs(a * ix) + np.abs(b * iy) + abs(c)
            sy = np.abs(d * ix) + np.abs(e * iy) + abs(f)
            px, py = px + offset[0] * sx, py + offset[1] * sy
        rx, ry = np.zeros_like(px), np.zeros_like(py)
        for slot in vs["vars"]:
            prm = {pn: float(P[off]) for pn, off in slot["params"].items()}
            vx, vy = fe.evaluate(mdb[slot["name"]], commons, px, py, float(P[slot["weight"]]), prm, aff, rng=rng,
                                 with_precalc=False, mca=mca)
            rx, ry = rx + vx, ry + vy
        ix, iy = rx, ry
    return ix, iy, sx, sy


def step_reference(mdb, commons, L, P, xi, x, y, words, precision, seed=0, input_ulp=None):
    """float64 step of xform xi for walkers at (x, y) with their words:
    (rx, ry, spread_x, spread_y, mca_x, mca_y, any_bad, any_ok); input_ulp:
    the relative rounding of the input point (default: the kernel's)"""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    with np.errstate(all="ignore"):
        rx, ry, sx, sy = _xform_eval(mdb, commons, L, P, xi, x, y, words, precision)
        if input_ulp is None:
            input_ulp = ULP if precision == 0 else 2.0 ** -52
        ulp_in = 4 * input_ulp
        spx, spy = np.zeros_like(rx), np.zeros_like(ry)
        for off in ((ulp_in, 0.0), (-ulp_in, 0.0), (0.0, ulp_in), (0.0, -ulp_in)):
            qx, qy, _, _ = _xform_eval(mdb, commons, L, P, xi, x, y, words, precision, offset=off)
            spx, spy = np.maximum(spx, jump(qx, rx)), np.maximum(spy, jump(qy, ry))
        pert = (dict(op_rel=2.0 ** -24, fn_rel=2.0 ** -21, fn_abs=2.0 ** -21) if precision == 0 else
                dict(op_rel=2.0 ** -53, fn_rel=2.0 ** -51, fn_abs=2.0 ** -51))
        mrng = np.random.default_rng(seed)
        nx, ny = np.zeros_like(rx), np.zeros_like(ry)
        any_bad = np.zeros(rx.shape, dtype=bool)
        any_ok = np.zeros(rx.shape, dtype=bool)
        for _ in range(MCA_RUNS):
            qx, qy, _, _ = _xform_eval(mdb, commons, L, P, xi, x, y, words, precision, mca=fe.Perturb(mrng, **pert))
            nx, ny = np.maximum(nx, jump(qx, rx)), np.maximum(ny, jump(qy, ry))
            ok = (np.abs(qx) <= 1e9) & (np.abs(qy) <= 1e9)
            any_ok |= ok
            any_bad |= ~ok
    return rx, ry, spx, spy, nx, ny, any_bad, any_ok


def check_steps(gpu_x, gpu_y, gpu_op, ref, K) -> dict:
    """the bound over a set of steps; returns counts and the largest K needed"""
    rx, ry, sx, sy, nx, ny, any_bad, any_ok = ref
    gx, gy = np.asarray(gpu_x, np.float64), np.asarray(gpu_y, np.float64)
    reset = np.asarray(gpu_op) == 0.0
    with np.errstate(invalid="ignore", over="ignore"):
        big = np.maximum(1.0, np.maximum(np.abs(rx), np.abs(ry)))
        ref_ok = (np.abs(rx) <= 1e9) & (np.abs(ry) <= 1e9)
        knife = ~(np.maximum(np.maximum(sx, sy), np.maximum(nx, ny)) <= 1e-3 * big)
        tol_x = K * ULP * np.maximum(1.0, np.abs(rx)) + 2.0 * sx + 2.0 * nx
        tol_y = K * ULP * np.maximum(1.0, np.abs(ry)) + 2.0 * sy + 2.0 * ny
        ex, ey = np.abs(gx - rx), np.abs(gy - ry)
        checked = ~reset & ref_ok & ~knife
        bad = checked & ~((ex <= tol_x) & (ey <= tol_y))
        need = np.where(checked, np.maximum((ex - 2 * sx - 2 * nx) / np.maximum(1.0, np.abs(rx)),
                                            (ey - 2 * sy - 2 * ny) / np.maximum(1.0, np.abs(ry))) / ULP, -np.inf)
        rel_tol = np.maximum(tol_x / np.maximum(1.0, np.abs(rx)), tol_y / np.maximum(1.0, np.abs(ry)))
    return {"steps": int(gx.size), "checked": int(checked.sum()), "knife": float(knife.mean()),
            "bad": int(bad.sum()), "wrong_reset": int((reset & ~any_bad & ~knife).sum()),
            "missed_reset": int((~reset & ~ref_ok & ~any_ok & ~knife).sum()),
            "k_needed": float(need.max()) if need.size else 0.0,
            "median_rel_bound": float(np.median(rel_tol[checked])) if checked.any() else 0.0,
            "first_bad": None if not bad.any() else int(np.nonzero(bad)[0][0])}


def trace_bound(mdb, commons, L, P_blocks, tpb, seed, stream, pass0, x0, tr, xid, precision, K,
                per_walker_xid=False) -> dict:
    """the bound over every pass of a one-step trace: pass i's input is the
    GPU's output of pass i-1 after the block permutation (walkers keep their
    point in xaos flames, per_walker_xid, where the pass's first lane word
    picks the xform), pass 0's the state x0"""
    npass, nw = tr.shape[0], tr.shape[1]
    bits = int(np.log2(tpb))
    blocks = nw // tpb
    gids = np.arange(nw, dtype=np.uint64)
    xin = np.asarray(x0[:, 0], np.float64).copy()
    yin = np.asarray(x0[:, 1], np.float64).copy()
    shared = bool(np.all(P_blocks[:, 12:] == P_blocks[0, 12:]))
    totals = {"steps": 0, "checked": 0, "bad": 0, "wrong_reset": 0, "missed_reset": 0, "k_needed": -np.inf,
              "knife_steps": 0, "median_rel_bound": []}
    for i in range(npass):
        pass_ = pass0 + i
        if i > 0:
            px, py = tr[i - 1][:, 0].astype(np.float64), tr[i - 1][:, 1].astype(np.float64)
            if per_walker_xid:
                xin, yin = px, py
            else:
                # slot perm(t) of each block receives walker t's point
                key = oracle.philox(np.array([[(pass_ - 1) >> 1, b, 0, oracle.TAG_PERM] for b in range(blocks)],
                                             dtype=np.uint32), seed, stream)
                xin, yin = np.empty(nw), np.empty(nw)
                t = np.arange(tpb, dtype=np.uint64)
                for b in range(blocks):
                    odd = (pass_ - 1) & 1
                    wa, wb = int(key[b][2 if odd else 0]), int(key[b][3 if odd else 1])
                    slot = _permute(t, wa, wb, bits)
                    xin[b * tpb + slot] = px[b * tpb:(b + 1) * tpb]
                    yin[b * tpb + slot] = py[b * tpb:(b + 1) * tpb]
        words = lane_words(gids, seed, stream, pass_)
        if per_walker_xid:
            words = words[:, 1:]   # the pass's first draw picked the xform
        xw = xid[i] if per_walker_xid else xid[i][np.arange(nw) // 64]
        for xi in np.unique(xw):
            m = np.nonzero(xw == xi)[0]
            blk = m // tpb
            # blocks of a static flame share every slot past the screen and
            # plane affines (only the screen's anti-alias jitter differs)
            groups = [(0, m)] if shared else [(b, m[blk == b]) for b in np.unique(blk)]
            for b, mm in groups:
                # the f64 kernel carries its walkers in double between passes;
                # fed back from the float32 trace, they are known to float32
                iu = ULP if (precision == 1 and i > 0) else None
                ref = step_reference(mdb, commons, L, P_blocks[b].astype(np.float64), int(xi), xin[mm], yin[mm],
                                     words[mm], precision, seed=1000 * i + int(xi), input_ulp=iu)
                r = check_steps(tr[i][mm, 0], tr[i][mm, 1], tr[i][mm, 3], ref, K)
                for k in ("steps", "checked", "bad", "wrong_reset", "missed_reset"):
                    totals[k] += r[k]
                totals["knife_steps"] += int(round(r["knife"] * r["steps"]))
                totals["k_needed"] = max(totals["k_needed"], r["k_needed"])
                if r["checked"]:
                    totals["median_rel_bound"].append(r["median_rel_bound"])
    totals["median_rel_bound"] = float(np.median(totals["median_rel_bound"]))
    return totals


def _permute(t, wa, wb, bits):
    """the block permutation (flame_device.h rfk_permute)"""
    mask = np.uint64((1 << bits) - 1)
    sh = np.uint64(bits >> 1 if bits > 1 else 1)
    x = (t * np.uint64((wa & 0xFFFF) | 1) + np.uint64(wa >> 16)) & mask
    x ^= x >> sh
    x = (x * np.uint64((wb & 0xFFFF) | 1) + np.uint64(wb >> 16)) & mask
    x ^= x >> sh
    return x.astype(np.int64)

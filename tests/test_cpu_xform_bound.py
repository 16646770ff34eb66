"""The one-step bound of tests/xform_bound.py on the oracle alone: the
oracle's own trace (double, stored as float32 as the GPU stores it) meets the
bound, and traces with a planted mistake do not."""
import numpy as np
import pytest

import xform_bound as xb
from oracle import oracle
from refraktor_amd import flame as fl
from refraktor_amd import render

W, H = 320, 180
BLOCKS, TPB = 16, 256


@pytest.fixture(scope="module", params=["small", "1080"])
def traced(request):
    f = fl.flame_small(3) if request.param == "small" else fl.flame_1080(3)
    L = oracle.Layout(f, oracle.load_variation_params(render.CONFIG_DIR))
    _, P, _ = oracle.prepare(f, render.CONFIG_DIR, W, H, BLOCKS, 9)
    x, y, c = oracle.seed_walkers(P, P.shape[1], BLOCKS, TPB, W, H, 9)
    x0 = np.stack([x, y, c], 1).astype(np.float32)
    tr, xid = oracle.trace(L, P, P.shape[1], BLOCKS, TPB, 9, 3, 6, x0[:, 0], x0[:, 1], x0[:, 2])
    mdb, cm = xb.load_db(render.CONFIG_DIR)
    return L, P, x0, tr.astype(np.float32), xid, mdb, cm


def _bound(traced, tr, xid):
    L, P, x0, _, _, mdb, cm = traced
    return xb.trace_bound(mdb, cm, L, P, TPB, 9, 0, 3, x0, tr, xid, 0, 8)


def test_oracle_trace_meets_the_bound(traced):
    r = _bound(traced, traced[3], traced[4])
    assert r["bad"] == 0 and r["wrong_reset"] == 0 and r["missed_reset"] == 0, r
    assert r["checked"] > 0.98 * r["steps"]
    assert r["median_rel_bound"] < 1e-5


def test_planted_output_error_fails(traced):
    tr, xid = traced[3].copy(), traced[4]
    wave_x = xid[2][np.arange(tr.shape[1]) // 64]
    m = wave_x == wave_x[0]
    tr[2][m, 0] *= np.float32(1.0 + 2e-5)   # one xform's x output off by 2e-5
    assert _bound(traced, tr, xid)["bad"] > 0


def test_planted_xform_choice_fails(traced):
    xid = traced[4].copy()
    xid[1] = (xid[1] + 1) % traced[0].n_xforms
    assert _bound(traced, traced[3], xid)["bad"] > 0

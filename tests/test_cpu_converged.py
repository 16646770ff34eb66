"""The converged-image statistic (tests/parity_stats.py) on the oracle alone:
independent seeds stay within TAU, and mistakes planted in the oracle's
parameters (in its lightest xform) exceed it. 640x360 C1 flame of seed 3,
2^16 walkers x 1024 passes + 20 fuse, as the GPU test renders it."""
import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import flame as fl
from refraktor_amd import render
from parity_stats import TAU, converged_ratio

W, H = 640, 360
BLOCKS, TPB, PASSES, FUSE = 256, 256, 1024, 20


@pytest.fixture(scope="module")
def c1():
    f = fl.flame_small(3)
    L = oracle.Layout(f, oracle.load_variation_params(render.CONFIG_DIR))
    _, P, pal = oracle.prepare(f, render.CONFIG_DIR, W, H, BLOCKS, 13)

    def run(params, seed):
        h, _ = oracle.render(L, params, params.shape[1], pal, BLOCKS, TPB, W, H, seed, FUSE, PASSES)
        return h

    ref, other = run(P, 14), run(P, 15)
    return f, L, P, run, ref, other


def test_independent_seeds_within_tau(c1):
    _, _, P, run, ref, other = c1
    ratio, s, s0, n = converged_ratio(run(P, 13), ref, other)
    assert n > 10000
    assert ratio <= TAU, (ratio, s, s0)


@pytest.mark.parametrize("what", ["affine", "variation_weight", "colour"])
def test_planted_mistakes_exceed_tau(c1, what):
    f, L, P, run, ref, other = c1
    light = int(np.argmin([x["weight"] for x in f["xforms"]]))
    xs = L.xforms[light]
    bad = P.copy()
    if what == "affine":
        bad[:, xs["vlinks"][0]["aff"]] *= 1.005
    elif what == "variation_weight":
        bad[:, xs["vlinks"][0]["vars"][0]["weight"]] *= 1.02
    else:
        bad[:, xs["color"]] = bad[:, xs["color"]] * 1.03 + 0.01
    ratio, s, s0, _ = converged_ratio(run(bad, 13), ref, other)
    assert ratio > TAU, (what, ratio, s, s0)

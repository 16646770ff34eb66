"""Animated parameters (anima::sample): `increase` and `sine` as the
reference's front end registers them (src/refraktor/src/refraktor.cpp:351-367)
and the built-in `mix` (src/librefrakt/src/anima.cpp:35-105). The oracle's
evaluation is pinned by hand-computed values; the library's packed samples of
an animated flame must equal the oracle's at every time."""
import math

import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import flame as fl
from refraktor_amd import render


def A(t0, call=None, **args):
    return {"t0": t0, "call": call, "args": args} if call else t0


@pytest.mark.parametrize("v,t,want", [
    (A(10.0, "increase", per_loop=90.0), 0.5, 55.0),
    (A(10.0, "increase"), 0.25, 100.0),                        # per_loop defaults to 360
    (A(0.2, "sine", amplitude=0.3), 0.25, 0.5),                 # sin(pi/2)
    (A(0.2, "sine", amplitude=0.3), 0.75, -0.1),
    (A(0.2, "sine", amplitude=0.3, phase=90.0), 0.0, 0.5),      # phase in degrees
    (A(0.2, "sine", amplitude=0.3, frequency=2.0), 0.125, 0.5),
    (A(0.2, "sine", amplitude=0.3, sharpness=2.0), 1.0 / 12.0, 0.275),        # sign(v) |v|^2, v = 1/2
    (A(0.2, "sine", amplitude=0.3, sharpness=2.0), 7.0 / 12.0, 0.125),        # v = -1/2
    (A(0.2, "sine", amplitude=0.3, absolute=True), 7.0 / 12.0, 0.35),
])
def test_increase_and_sine(v, t, want):
    assert abs(oracle._anima(v, t) - want) < 1e-12


def test_mix_smoothstep():
    mix = A(1.0, "mix", **{"right.value": 3.0, "start_time": 0.2, "length": 0.4})
    assert oracle._anima(mix, 0.1) == 1.0          # before: the left (initial) value
    assert oracle._anima(mix, 0.2) == 1.0          # at the start
    assert oracle._anima(mix, 0.6) == 3.0          # at the end
    assert oracle._anima(mix, 0.9) == 3.0
    assert abs(oracle._anima(mix, 0.4) - 2.0) < 1e-12        # quintic smoothstep(1/2) = 1/2
    # smoothstep(1/4) = 6/4^5 - 15/4^4 + 10/4^3 = 0.103515625
    assert abs(oracle._anima(mix, 0.3) - (1.0 + 2.0 * 0.103515625)) < 1e-12
    # nested: the left side increases, the right side is a sine of its own value
    nested = A(1.0, "mix", **{"left.function": "increase", "left.per_loop": 100.0, "right.value": 3.0,
                              "right.function": "sine", "right.amplitude": 0.5, "start_time": 0.2, "length": 0.4})
    assert abs(oracle._anima(nested, 0.1) - 11.0) < 1e-12
    assert abs(oracle._anima(nested, 0.75) - (3.0 + 0.5 * math.sin(1.5 * math.pi))) < 1e-12
    # missing pieces: the initial value, as the reference returns after its error log
    assert oracle._anima(A(1.0, "mix", **{"start_time": 0.0, "length": 1.0}), 0.5) == 1.0
    assert oracle._anima(A(1.0, "mix", **{"right.value": 2.0, "start_time": 0.0}), 0.5) == 1.0
    assert oracle._anima(A(4.0, "no_such_function"), 0.5) == 4.0


def animated_flame() -> dict:
    f = fl.flame_small(3)
    f["rotate"] = A(10.0, "increase", per_loop=90.0)
    f["scale"] = A(0.25, "sine", amplitude=0.05, frequency=0.5)
    x0, x1, x2 = f["xforms"]
    x0["weight"] = A(x0["weight"], "sine", amplitude=0.1, phase=30.0)
    x1["color"] = A(0.5, "mix", **{"right.value": 0.9, "start_time": 0.25, "length": 0.5})
    x2["vchain"][0]["mod_rotate"] = A(0.0, "increase", per_loop=45.0)
    x2["vchain"][0]["transform"][4] = A(x2["vchain"][0]["transform"][4], "sine", amplitude=0.2, sharpness=1.5)
    x1["vchain"][0]["variations"] = {"sinusoidal": A(1.0, "sine", amplitude=0.25, absolute=True),
                                     "curl": {"weight": 0.5, "parameters": {"c1": A(1.0, "increase", per_loop=-0.5),
                                                                            "c2": 0.1}}}
    return f


def test_library_pack_equals_oracle_pack():
    db = render.VariationDB()
    f = animated_flame()
    fo = db.flame(f)
    L = oracle.Layout(fo.to_json(), oracle.load_variation_params(render.CONFIG_DIR))
    for start, offset in ((0.0, 0.3), (0.7, 0.11), (-0.4, 0.5)):
        got = fo.pack_samples(320, 180, start, offset, 4)
        want = oracle.pack_samples(f, L, 320, 180, start, offset, 4)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # the animation is not a no-op: the samples differ over time
    s = fo.pack_samples(320, 180, 0.0, 0.3, 4).reshape(4, -1)
    assert not np.allclose(s[0], s[3])

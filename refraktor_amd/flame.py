"""Host-side flame descriptions and the seeded synthetic flame generator.

A flame is a plain dict in librefrakt's flame JSON shape
(src/librefrakt/src/flame_types.cpp:366-485): view and tone parameters,
`xforms` (weight, color, color_speed, opacity, vchain of vlinks with a
6-coefficient transform in (a, d, b, e, c, f) order and a variation map),
optional `final_xform`, and a 256-entry (h, s, v) palette.

The generator draws every random number from one SplitMix64 stream per seed,
in the order the functions below consume it.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass

MASK64 = (1 << 64) - 1

# variation set of the 1080p / 4K / animation workloads
CONFIG_VARIATIONS = ["julia", "swirl", "horseshoe", "polar", "handkerchief", "heart", "disc", "spiral",
                     "hyperbolic", "diamond", "ex", "bubble", "curl"]


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """double in [0, 1) from the top 53 bits"""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_pm1(self) -> float:
        return self.uniform() * 2.0 - 1.0

    def randint(self, lo: int, hi: int) -> int:
        """integer in [lo, hi]"""
        return lo + int(self.uniform() * (hi - lo + 1))


def rgb_to_hsv(r: float, g: float, b: float) -> tuple[float, float, float]:
    """librefrakt color::rgb_to_hsv (src/librefrakt/src/util/color.cpp:5-27)"""
    cmax, cmin = max(r, g, b), min(r, g, b)
    delta = cmax - cmin
    h, s = 0.0, 0.0
    if delta > 0.0:
        s = delta / cmax
        if r == cmax:
            h = 60.0 * math.fmod((g - b) / delta, 6.0)
        elif g == cmax:
            h = 60.0 * ((b - r) / delta + 2.0)
        else:
            h = 60.0 * ((r - g) / delta + 4.0)
    return (h, s, cmax)


def vlink(transform, variations: dict, mod_scale=1.0, mod_rotate=0.0, mod_x=0.0, mod_y=0.0) -> dict:
    return {"transform": list(transform), "mod_x": mod_x, "mod_y": mod_y, "mod_scale": mod_scale,
            "mod_rotate": mod_rotate, "variations": dict(variations)}


def xform(weight, color, color_speed, vchain: list, opacity=1.0) -> dict:
    return {"weight": weight, "color": color, "color_speed": color_speed, "opacity": opacity,
            "vchain": list(vchain)}


def make_flame(xforms: list, palette: list, final_xform: dict | None = None, *, center=(0.0, 0.0), scale=0.25,
               rotate=0.0, gamma=2.2, brightness=4.0, vibrancy=1.0) -> dict:
    f = {"center_x": center[0], "center_y": center[1], "scale": scale, "rotate": rotate, "gamma": gamma,
         "brightness": brightness, "vibrancy": vibrancy, "highlight_power": 1.0, "gamma_threshold": 0.0,
         "mod_hue": 0.0, "mod_sat": 0.0, "mod_val": 0.0, "xforms": xforms, "palette": palette}
    if final_xform is not None:
        f["final_xform"] = final_xform
    return f


def to_json(flame: dict) -> str:
    return json.dumps(flame)


# ------------------------------------------------------------- generator

def random_palette(rng: SplitMix64) -> list:
    pal = []
    for _ in range(256):
        r, g, b = (rng.randint(0, 255) for _ in range(3))
        pal.append(list(rgb_to_hsv(r / 255.0, g / 255.0, b / 255.0)))
    return pal


def random_affine(rng: SplitMix64, scale: float = 1.0) -> list:
    return [rng.uniform_pm1() * scale for _ in range(6)]


def dirichlet_weights(rng: SplitMix64, n: int) -> list:
    e = [-math.log(1.0 - rng.uniform()) for _ in range(n)]
    s = sum(e)
    return [v / s for v in e]


def variation_entry(name: str, weight: float):
    """variation JSON value; parameterised variations use their defaults"""
    if name == "curl":
        return {"weight": weight, "parameters": {"c1": 1.0, "c2": 0.0}}
    return weight


def _xforms(rng: SplitMix64, n: int, var_choice) -> list:
    weights = dirichlet_weights(rng, n)
    out = []
    for i in range(n):
        aff = random_affine(rng)
        vars_ = var_choice(rng, i)
        color = i / (n - 1) if n > 1 else 0.0
        out.append(xform(weights[i], color, 0.5, [vlink(aff, vars_)]))
    return out


def flame_small(seed: int) -> dict:
    """640x360 plain render: 3 xforms (linear, sinusoidal, spherical), weight 1 each
    variation, affine U[-1,1], xform weights Dirichlet(1), random palette."""
    rng = SplitMix64(seed)
    names = ["linear", "sinusoidal", "spherical"]
    xfs = _xforms(rng, 3, lambda r, i: {names[i]: 1.0})
    return make_flame(xfs, random_palette(rng), gamma=2.2, brightness=4.0, vibrancy=1.0)


def _multi_variation(max_vars: int, min_vars: int = 2):
    def choose(rng: SplitMix64, i: int) -> dict:
        k = rng.randint(min_vars, max_vars)
        pool = list(CONFIG_VARIATIONS)
        chosen = {}
        for _ in range(k):
            j = rng.randint(0, len(pool) - 1)
            name = pool.pop(j)
            chosen[name] = variation_entry(name, rng.uniform())
        return chosen
    return choose


def flame_1080(seed: int) -> dict:
    """1920x1080 plain render: 8 xforms of 2-4 variations (weights U[0,1]) from the
    13-variation set, plus a final xform (linear, affine U[-1,1] x 0.5)."""
    rng = SplitMix64(seed)
    xfs = _xforms(rng, 8, _multi_variation(4, 2))
    final = xform(0.0, 0.5, 0.0, [vlink(random_affine(rng, 0.5), {"linear": 1.0})])
    return make_flame(xfs, random_palette(rng), final, vibrancy=1.0)


def flame_4k(seed: int) -> dict:
    """3840x2160 plain render: 12 xforms of 1-4 variations."""
    rng = SplitMix64(seed)
    xfs = _xforms(rng, 12, _multi_variation(4, 1))
    final = xform(0.0, 0.5, 0.0, [vlink(random_affine(rng, 0.5), {"linear": 1.0})])
    return make_flame(xfs, random_palette(rng), final, vibrancy=1.0)


def flame_keyframe(seed: int) -> dict:
    """animation keyframe: 6-10 xforms of 2-4 variations, no final xform"""
    rng = SplitMix64(seed)
    n = rng.randint(6, 10)
    xfs = _xforms(rng, n, _multi_variation(4, 2))
    return make_flame(xfs, random_palette(rng), vibrancy=1.0)


def flame_xaos(seed: int) -> dict:
    """5 xforms of 1-2 variations with a xaos matrix: entries U[0,2], a third
    of them zero, no all-zero row"""
    rng = SplitMix64(seed)
    n = 5
    xfs = _xforms(rng, n, _multi_variation(2, 1))
    chaos = []
    for i in range(n):
        row = [0.0 if rng.uniform() < 1.0 / 3.0 else 2.0 * rng.uniform() for _ in range(n)]
        if not any(row):
            row[i] = 1.0
        chaos.append(row)
    f = make_flame(xfs, random_palette(rng), vibrancy=1.0)
    f["chaos"] = chaos
    return f


@dataclass
class Workload:
    name: str
    width: int
    height: int
    iterations: int
    walkers: int
    fuse: int
    de_radius: float


WORKLOADS = {
    "small": Workload("plain render 640x360, 3 xforms, 2^26 iterations, 2^16 walkers", 640, 360, 1 << 26,
                      1 << 16, 20, 0.0),
    "1080p": Workload("plain render 1920x1080, 8 xforms + final, 2^30 iterations, DE radius 9", 1920, 1080, 1 << 30,
                      1 << 20, 20, 9.0),
    "4k": Workload("plain render 3840x2160, 12 xforms + final, 2^32 iterations, DE radius 12", 3840, 2160, 1 << 32,
                   1 << 21, 20, 12.0),
}

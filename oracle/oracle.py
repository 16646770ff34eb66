"""ORACLE — test infrastructure only (see flame_oracle.c for the C half).

Independent restatements, in numpy / plain Python, of the pieces of the path
that are not in the C file:
  * the packed flame layout and pack_sample (flame_types.h:602-712), computed
    from the flame dict and the variation parameter lists read with PyYAML;
  * fix_rotation (flame_compiler.cpp:957-1035), Catmull-Rom coefficients
    (catmull.cu:11-42), per-block sample evaluation with polar affines and the
    HSV palette (refactor.cu:105-151, 256-294);
  * precalc: weight sum, anti-alias jitter, the variations' precalc blocks;
  * tonemap (tonemap.cu:50-105, tonemapper.cpp:52), the log2 density
    histogram (density_histo.cu:1-41), and a brute-force density estimation;
  * Philox4x32-10 vectorised in numpy.
Only tests/, the smoke check and bench.py's cpu_baseline may use this module.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np
import yaml

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

# variation ids of flame_oracle.c
VARIATIONS = ["linear", "sinusoidal", "spherical", "swirl", "horseshoe", "polar", "handkerchief", "heart", "disc",
              "spiral", "hyperbolic", "diamond", "ex", "julia", "bubble", "curl", "fisheye", "eyefish", "blur",
              "gaussian_blur", "waves", "julian", "bent", "cylinder", "exponential", "power", "cosine", "noise",
              "rectangles", "pdj", "cross", "exp", "log", "sinh", "cosh", "sin", "cos", "tangent", "square",
              "polar2", "foci", "perspective", "blob", "rings", "fan", "bent2", "ngon"]
VAR_ID = {n: i for i, n in enumerate(VARIATIONS)}

TAG_PICK, TAG_PERM, TAG_LANE, TAG_INIT, TAG_JITTER = 1, 2, 3, 4, 5


# ------------------------------------------------------------------ Philox
def philox(ctr: np.ndarray, key0: int, key1: int) -> np.ndarray:
    """Philox4x32-10 over rows of a (N, 4) uint32 counter array."""
    c = np.array(ctr, dtype=np.uint64).reshape(-1, 4) & 0xFFFFFFFF
    k0, k1 = np.uint64(key0 & 0xFFFFFFFF), np.uint64(key1 & 0xFFFFFFFF)
    M = np.uint64(0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(0x9E3779B9)) & M
            k1 = (k1 + np.uint64(0xBB67AE85)) & M
        p0 = np.uint64(0xD2511F53) * c[:, 0]
        p1 = np.uint64(0xCD9E8D57) * c[:, 2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & M
        hi1, lo1 = p1 >> np.uint64(32), p1 & M
        c = np.stack([hi1 ^ c[:, 1] ^ k0, lo1, hi0 ^ c[:, 3] ^ k1, lo0], axis=1)
    return c.astype(np.uint32)


def u01(u) -> np.ndarray:
    return np.asarray(u, dtype=np.uint32).astype(np.float32) * np.float32(2.3283064365386963e-10)


# --------------------------------------------------------------- layout
def load_variation_params(config_dir: str | Path) -> dict:
    """name -> (sorted user params, sorted precalc params, defaults)"""
    with open(Path(config_dir) / "variations_new.yml") as fh:
        y = yaml.safe_load(fh)
    out = {}
    for name, v in y.items():
        user, pre, dflt = [], [], {}
        if isinstance(v, dict):
            for pn, pd in (v.get("param") or {}).items():
                pd = pd or {}
                if "precalc" in (pd.get("tags") or []):
                    pre.append(pn)
                else:
                    user.append(pn)
                    dflt[pn] = float(pd.get("default", 0.0))
        out[name] = (sorted(user), sorted(pre), dflt)
    return out


def _anima(v, t=0.0) -> float:
    """an animated value at time t (anima::sample): its initial value t0,
    through its call when it has one"""
    if not isinstance(v, dict):
        return float(v)
    iv = float(v["t0"])
    name = v.get("call")
    return _anima_call(name, t, iv, v.get("args", {}), "") if name else iv


def _anima_call(name, t, iv, args, prefix) -> float:
    """`increase` and `sine` as the reference's front end registers them
    (src/refraktor/src/refraktor.cpp:351-367), `mix` the built-in quintic
    smoothstep cross-fade (src/librefrakt/src/anima.cpp:35-105); anything
    missing or unknown gives the initial value, as the reference does after
    logging an error"""
    def num(k, d=None):
        v = args.get(prefix + k, d)
        return float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else d
    if name == "increase":
        return iv + t * num("per_loop", 360.0)
    if name == "sine":
        v = math.sin(t * num("frequency", 1.0) * math.pi * 2.0 + math.radians(num("phase", 0.0)))
        sharp = num("sharpness", 0.0)
        if sharp > 0:
            v = math.copysign(1.0, v) * abs(v) ** sharp
        if args.get(prefix + "absolute", False) is True:
            v = abs(v)
        return iv + v * num("amplitude", 1.0)
    if name == "mix":
        left = iv
        lf = args.get(prefix + "left.function")
        if isinstance(lf, str):
            left = _anima_call(lf, t, iv, args, prefix + "left.")
        right = num("right.value")
        if right is None:
            return iv
        rf = args.get(prefix + "right.function")
        if isinstance(rf, str):
            right = _anima_call(rf, t, right, args, prefix + "right.")
        start, length = num("start_time"), num("length")
        if start is None or length is None:
            return iv
        if t <= start:
            return left
        if t >= start + length:
            return right
        m = (t - start) / length
        m = 6 * m ** 5 - 15 * m ** 4 + 10 * m ** 3
        return left * (1 - m) + right * m
    return iv


class Layout:
    """Offsets of every slot in the packed parameter array."""

    def __init__(self, flame: dict, vparams: dict):
        self.flame = flame
        self.vparams = vparams
        n = len(flame["xforms"])
        self.n_xforms = n
        self.has_final = "final_xform" in flame
        self.chaos = "chaos" in flame
        idx = 13
        self.chaos_base = -1
        if self.chaos:
            self.chaos_base = idx
            idx += n * (n + 1)
        self.affine_indices = [0, 6]
        self.xforms = []
        for xf in flame["xforms"] + ([flame["final_xform"]] if self.has_final else []):
            xs = {"weight": idx, "color": idx + 1, "color_speed": idx + 2, "opacity": idx + 3, "vlinks": []}
            idx += 4
            for vl in xf["vchain"]:
                vs = {"aff": idx, "vars": []}
                self.affine_indices.append(idx)
                idx += 6
                for name in sorted(vl["variations"]):
                    user, pre, _ = vparams[name]
                    slot = {"name": name, "weight": idx, "params": {}}
                    idx += 1
                    for pn in user:
                        slot["params"][pn] = idx
                        idx += 1
                    for pn in pre:
                        slot["params"][pn] = idx
                        idx += 1
                    vs["vars"].append(slot)
                xs["vlinks"].append(vs)
            self.xforms.append(xs)
        self.size_reals = idx

    def program(self) -> np.ndarray:
        """int program for flame_oracle.c"""
        prog = []
        for xs in self.xforms:
            prog += [len(xs["vlinks"]), xs["weight"], xs["color"], xs["color_speed"], xs["opacity"]]
            for vs in xs["vlinks"]:
                prog += [vs["aff"], len(vs["vars"])]
                for v in vs["vars"]:
                    if v["name"] not in VAR_ID:
                        raise KeyError(f"oracle has no variation {v['name']}")
                    names = sorted(v["params"])
                    prog += [VAR_ID[v["name"]], v["weight"], len(names)] + [v["params"][p] for p in names]
        return np.array(prog, dtype=np.int32)


def _rotated(a, deg):
    rad = -deg * math.pi / 180.0
    cs, sn = math.cos(rad), math.sin(rad)
    A, D, B, E, Cc, F = a
    return [A * cs + D * sn, -A * sn + D * cs, B * cs + E * sn, -B * sn + E * cs, Cc, F]


def _scaled(a, s):
    return [a[0] * s, a[1] * s, a[2] * s, a[3] * s, a[4], a[5]]


def _translated(a, x, y):
    return [a[0], a[1], a[2], a[3], a[0] * x + a[2] * y + a[4], a[1] * x + a[3] * y + a[5]]


def screen_space(flame, w, h, t=0.0):
    aff = [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    aff = _translated(aff, w / 2.0, h / 2.0)
    aff = _scaled(aff, _anima(flame["scale"], t) * h)
    aff = _rotated(aff, _anima(flame["rotate"], t))
    return _translated(aff, -_anima(flame["center_x"], t), -_anima(flame["center_y"], t))


def plane_space(flame, w, h, t=0.0):
    aff = [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]
    aff = _translated(aff, _anima(flame["center_x"], t), _anima(flame["center_y"], t))
    aff = _rotated(aff, -_anima(flame["rotate"], t))
    aff = _scaled(aff, 1.0 / (_anima(flame["scale"], t) * h))
    return _translated(aff, w / -2.0, h / -2.0)


def pack_samples(flame: dict, L: Layout, w: int, h: int, start: float, offset: float, count: int) -> np.ndarray:
    """flame::pack_samples (flame_types.h:707-712): samples at start + offset * i"""
    return np.concatenate([pack_sample(flame, L, w, h, start + offset * i) for i in range(count)])


def pack_sample(flame: dict, L: Layout, w: int, h: int, t: float = 0.0) -> np.ndarray:
    """one packed sample at time t (flame_types.h:634-672)"""
    out = np.zeros(L.size_reals + 768)
    out[0:6] = screen_space(flame, w, h, t)
    out[6:12] = plane_space(flame, w, h, t)
    if L.chaos:
        n = L.n_xforms
        for i, row in enumerate(flame["chaos"]):
            base = L.chaos_base + i * (n + 1)
            for j, v in enumerate(row):
                out[base + 1 + j] = _anima(v, t)
    xfs = flame["xforms"] + ([flame["final_xform"]] if L.has_final else [])
    for xf, xs in zip(xfs, L.xforms):
        out[xs["weight"]] = _anima(xf["weight"], t)
        out[xs["color"]] = _anima(xf["color"], t)
        out[xs["color_speed"]] = _anima(xf["color_speed"], t)
        out[xs["opacity"]] = _anima(xf["opacity"], t)
        for vl, vs in zip(xf["vchain"], xs["vlinks"]):
            a = [_anima(v, t) for v in vl["transform"]]
            a = _scaled(a, _anima(vl["mod_scale"], t))
            a = _rotated(a, _anima(vl["mod_rotate"], t))
            a = _translated(a, _anima(vl["mod_x"], t), _anima(vl["mod_y"], t))
            out[vs["aff"]:vs["aff"] + 6] = a
            for slot in vs["vars"]:
                val = vl["variations"][slot["name"]]
                user, pre, dflt = L.vparams[slot["name"]]
                if isinstance(val, dict) and "weight" in val:
                    out[slot["weight"]] = _anima(val["weight"], t)
                    given = val.get("parameters", {})
                else:
                    out[slot["weight"]] = _anima(val, t)
                    given = {}
                for pn in user:
                    out[slot["params"][pn]] = _anima(given.get(pn, dflt.get(pn, 0.0)), t)
    base = L.size_reals
    for j, (hh, ss, vv) in enumerate(flame["palette"]):
        out[base + 3 * j] = math.fmod(hh + _anima(flame["mod_hue"], t), 360.0)
        out[base + 3 * j + 1] = min(max(ss + _anima(flame["mod_sat"], t), 0.0), 1.0)
        out[base + 3 * j + 2] = min(max(vv + _anima(flame["mod_val"], t), 0.0), 1.0)
    return out


def fix_rotation(s: np.ndarray, sample_size: int, flame_size: int, indices, count: int) -> np.ndarray:
    s = s.copy()
    eps = 1e-10
    for idx in indices:
        for i in range(count):
            o = idx + i * sample_size
            ax, mx = math.atan2(s[o + 1], s[o]), math.hypot(s[o], s[o + 1])
            ay, my = math.atan2(s[o + 3], s[o + 2]), math.hypot(s[o + 2], s[o + 3])
            if mx == 0.0 and my != 0.0:
                ax = ay
            if my == 0.0 and mx != 0.0:
                ay = ax
            s[o], s[o + 1] = ax, ay
            s[o + 2] = math.log(mx) if mx > 0 else -700.0
            s[o + 3] = math.log(my) if my > 0 else -700.0
        for i in range(1, count):
            for j in range(2):
                me, prev = idx + i * sample_size + j, idx + (i - 1) * sample_size + j
                d = s[me] - s[prev]
                if d > math.pi + eps:
                    s[me] -= 2 * math.pi
                elif d < -(math.pi - eps):
                    s[me] += 2 * math.pi
    for i in range(flame_size, sample_size, 3):
        for j in range(1, count):
            me, prev = i + j * sample_size, i + (j - 1) * sample_size
            d = s[me] - s[prev]
            if d > 180.0 + eps:
                s[me] -= 360.0
            elif d < -(180.0 - eps):
                s[me] += 360.0
    return s


def catmull(samples: np.ndarray, sample_size: int) -> np.ndarray:
    """segments [nseg][sample_size][4] (a, b, c, d), tension 0, alpha 0.5"""
    s = samples.reshape(-1, sample_size)
    nseg = s.shape[0] - 3
    p0, p1, p2, p3 = s[0:nseg], s[1:nseg + 1], s[2:nseg + 2], s[3:nseg + 3]
    alpha = 0.5
    t01 = np.power(np.hypot(p1 - p0, 1.0), alpha)
    t12 = np.power(np.hypot(p2 - p1, 1.0), alpha)
    t23 = np.power(np.hypot(p3 - p2, 1.0), alpha)
    m1 = p2 - p1 + t12 * ((p1 - p0) / t01 - (p2 - p0) / (t01 + t12))
    m2 = p2 - p1 + t12 * ((p3 - p2) / t23 - (p3 - p1) / (t12 + t23))
    a = 2.0 * (p1 - p2) + m1 + m2
    b = -3.0 * (p1 - p2) - m1 - m1 - m2
    return np.stack([a, b, m1, p1], axis=-1)


def hsv_to_rgb8(h, s, v) -> np.ndarray:
    def f(n):
        k = np.fmod(n + h / 60.0, 6.0)
        return v - v * s * np.maximum(0.0, np.minimum(k, np.minimum(4.0 - k, 1.0)))

    def to8(c):
        return np.clip(np.floor(c * 255.0), 0, 255).astype(np.uint32)
    return to8(f(5.0)) | (to8(f(3.0)) << 8) | (to8(f(1.0)) << 16) | np.uint32(255 << 24)


def eval_samples(segments: np.ndarray, L: Layout, nsamples: int):
    """per-block params (double, before precalc) and RGBA8 palettes"""
    nseg = segments.shape[0]
    seg_size = max(nsamples // nseg, 1)
    out = np.zeros((nsamples, L.size_reals))
    pals = np.zeros((nsamples, 256), dtype=np.uint32)
    for s in range(nsamples):
        seg = min(s // seg_size, nseg - 1)
        t = (s % seg_size) / float(max(nsamples // nseg, 1))
        cf = segments[seg]
        vals = cf[:, 0] * t * t * t + cf[:, 1] * t * t + cf[:, 2] * t + cf[:, 3]
        p = vals[:L.size_reals].copy()
        for a in L.affine_indices:
            angx, angy, lmx, lmy = vals[a], vals[a + 1], vals[a + 2], vals[a + 3]
            mx, my = math.exp(lmx), math.exp(lmy)
            p[a], p[a + 1], p[a + 2], p[a + 3] = mx * math.cos(angx), mx * math.sin(angx), my * math.cos(angy), \
                my * math.sin(angy)
        out[s] = p
        pv = vals[L.size_reals:].reshape(256, 3)
        pals[s] = hsv_to_rgb8(pv[:, 0], pv[:, 1], pv[:, 2])
    return out, pals


def jitter(seed: int, stream: int, nsamples: int) -> np.ndarray:
    """anti-alias jitter (randgauss(1/3)) of each sample from its JITTER stream"""
    ctr = np.zeros((nsamples * 2, 4), dtype=np.uint32)
    ctr[:, 0] = np.repeat(np.arange(nsamples), 2)
    ctr[:, 2] = np.tile([0, 1], nsamples)
    ctr[:, 3] = TAG_JITTER
    w = philox(ctr, seed, stream).reshape(nsamples, 8)
    ang = (u01(w[:, 0]) * np.float32(2.0)).astype(np.float64)
    s = u01(w[:, 1:5]).astype(np.float64).sum(axis=1)
    r = (1.0 / 3.0) * (s - 2.0)
    return np.stack([r * np.cos(np.pi * ang), r * np.sin(np.pi * ang)], axis=1)


PRECALC = {
    "waves": lambda P, s, aff: {"dx2": 1.0 / (P[aff + 4] ** 2 or 1e-20), "dy2": 1.0 / (P[aff + 5] ** 2 or 1e-20)},
    "julian": lambda P, s, aff: {"rN": abs(P[s["params"]["power"]]),
                                 "cn": 0.5 * P[s["params"]["dist"]] / P[s["params"]["power"]]},
    "perspective": lambda P, s, aff: {"vsin": math.sin(math.pi * 0.5 * P[s["params"]["angle"]]),
                                      "vfcos": P[s["params"]["dist"]] * math.cos(math.pi * 0.5 * P[s["params"]["angle"]])},
}


def precalc(params: np.ndarray, L: Layout, seed: int, stream: int) -> np.ndarray:
    """weight sum, xaos rows, jitter, per-variation precalc (in place on a copy)"""
    P = params.copy()
    n = L.n_xforms
    P[:, 12] = sum(P[:, L.xforms[i]["weight"]] for i in range(n))
    if L.chaos:
        for i in range(n):
            row = L.chaos_base + i * (n + 1)
            for j in range(n):
                P[:, row + 1 + j] *= P[:, L.xforms[j]["weight"]]
                P[:, row] += P[:, row + 1 + j]
    J = jitter(seed, stream, P.shape[0])
    P[:, 4] += J[:, 0]
    P[:, 5] += J[:, 1]
    for b in range(P.shape[0]):
        for xs in L.xforms:
            for vs in xs["vlinks"]:
                for slot in vs["vars"]:
                    fn = PRECALC.get(slot["name"])
                    if fn:
                        for k, v in fn(P[b], slot, vs["aff"]).items():
                            P[b, slot["params"][k]] = v
    return P


def param_stride(size_reals: int, precision: int = 0) -> int:
    align = 8 if precision else 16
    return (size_reals + align - 1) // align * align


# ------------------------------------------------------- image functions
def tonemap(hist: np.ndarray, quality: float, gamma: float, brightness: float, vibrancy: float) -> np.ndarray:
    """tonemap.cu:50-105 in double; hist (...,4) u32 fixed point or float draw units"""
    col = hist.astype(np.float64)
    if hist.dtype == np.uint32:
        col = col / 255.0
    k = 10.0 ** (-math.log10(quality) - 0.5)
    w = col[..., 3]
    out = np.zeros(col.shape)
    nz = w > 0
    c = col[nz]
    wn = c[:, 3]
    factor = 0.5 * brightness * np.log10(1.0 + wn * k) / wn
    c = c * factor[:, None]
    inv_g = 1.0 / gamma
    gf = np.power(c[:, 3], inv_g) / c[:, 3]
    rgb = c[:, :3]
    res = np.clip(np.power(rgb, inv_g) + vibrancy * (rgb * gf[:, None] - np.power(rgb, inv_g)), 0.0, 1.0)
    alpha = np.clip(np.log2(wn) / 24.0, 0.0, 1.0)
    out[nz, :3] = res
    out[nz, 3] = alpha
    return out


def density_histogram(hist: np.ndarray):
    v = hist[..., 3].astype(np.float32).ravel() * np.float32(1.0 / 255.0)
    bins = np.zeros(97, dtype=np.uint64)
    big = v >= 1.0
    b = (np.log(v[big].astype(np.float32)) * np.float32(4.0 / math.log(2.0))).astype(np.int64) + 1
    b = np.minimum(b, 96)
    np.add.at(bins, b, 1)
    bins[0] += np.count_nonzero(~big)
    mx = int(v[big].max()) if big.any() else 0
    return bins, mx


def de_levels(hist: np.ndarray, max_r: float, min_r: float, curve: float, max_level: int) -> np.ndarray:
    w = hist[..., 3]
    n = w.astype(np.float32) * np.float32(1.0 / 255.0)
    h = np.float32(max_r) * np.power(np.maximum(n, np.float32(1.0)), np.float32(-curve), dtype=np.float32)
    h = np.maximum(h, np.float32(min_r))
    lvl = np.minimum((h * np.float32(16.0) + np.float32(0.5)).astype(np.int64), max_level)
    return np.where(w == 0, -1, lvl)


def density_estimate(hist: np.ndarray, max_r: float, min_r: float, curve: float) -> np.ndarray:
    """brute-force scatter: every non-empty bin spreads its RGBA (draw units)
    with exp(-2 (r/h)^2), r <= h, normalised to sum 1 over the disc
    (h = 0: the bin keeps its value). Summed offset by offset: for each
    window offset, every source adds its level's normalised weight there."""
    H, W = hist.shape[:2]
    hmax = max(max_r, min_r)
    nlev = int(hmax * 16.0 + 0.5) + 1
    R = (nlev - 1 + 15) // 16
    lv = de_levels(hist, max_r, min_r, curve, nlev - 1)
    col = hist.astype(np.float64) / 255.0
    dy, dx = np.mgrid[-R:R + 1, -R:R + 1]
    d2 = (dx * dx + dy * dy).astype(np.float64).ravel()
    h2 = (np.arange(nlev) / 16.0) ** 2
    # kernel[level, offset], each row normalised over its disc
    with np.errstate(divide="ignore", invalid="ignore"):
        ker = np.where(d2[None, :] <= h2[:, None],
                       np.where(d2[None, :] == 0, 1.0, np.exp(-2.0 * d2[None, :] / np.maximum(h2[:, None], 1e-300))),
                       0.0)
    ker /= ker.sum(axis=1, keepdims=True)
    src = lv >= 0
    lvs = np.where(src, lv, 0)
    pad = np.zeros((H + 2 * R, W + 2 * R, 4))
    for j, (oy, ox) in enumerate(zip(dy.ravel(), dx.ravel())):
        wgt = np.where(src, ker[lvs, j], 0.0)
        if not wgt.any():
            continue
        pad[R + oy:R + oy + H, R + ox:R + ox + W] += wgt[..., None] * col
    return pad[R:R + H, R:R + W]


# ------------------------------------------------------------ C oracle
_clib = None


def clib():
    global _clib
    if _clib is None:
        srcs = [Path(__file__).with_name(n) for n in ("flame_oracle.c", "jpeg_oracle.c")]
        if not LIB.exists() or any(p.stat().st_mtime > LIB.stat().st_mtime for p in srcs):
            from . import build
            build.build()
        lib = C.CDLL(str(LIB))
        P = C.c_void_p
        lib.orc_philox.argtypes = [P, P, P]
        lib.orc_bin_points.argtypes = [P, P, P, C.c_size_t, P, C.c_uint32, C.c_uint32]
        lib.orc_trace.argtypes = [P, C.c_int, C.c_int, C.c_int, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P, P, P, P, P]
        lib.orc_trace.restype = C.c_int
        lib.orc_seed.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32, P, P, P]
        lib.orc_render.argtypes = [P, C.c_int, C.c_int, C.c_int, P, C.c_uint32, P, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_int]
        lib.orc_render.restype = C.c_uint64
        lib.orc_eval_variation.argtypes = [C.c_int, C.c_double, P, P, C.c_size_t, P, P, P, C.c_int, P, P]
        lib.orc_jpeg_encode.argtypes = [P, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_size_t]
        lib.orc_jpeg_encode.restype = C.c_long
        _clib = lib
    return _clib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def c_philox(ctr, key) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    clib().orc_philox(_ptr(ctr), _ptr(key), _ptr(out))
    return out


def eval_variation(name: str, weight: float, params: dict, aff: dict, px, py, words) -> tuple:
    """the C oracle's hand-written formula of one variation at points p
    (after the vlink affine); params holds every slot, user and precalc;
    words: per-point random words consumed in order"""
    prm = np.array([params[k] for k in sorted(params)] or [0.0], dtype=np.float64)
    a = np.array([aff[k] for k in "adbecf"], dtype=np.float64)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    w = np.ascontiguousarray(words, dtype=np.uint32)
    rx, ry = np.zeros_like(px), np.zeros_like(px)
    clib().orc_eval_variation(VARIATIONS.index(name), float(weight), _ptr(prm), _ptr(a), px.size, _ptr(px), _ptr(py),
                              _ptr(w), w.shape[1], _ptr(rx), _ptr(ry))
    return rx, ry


def bin_points(screen6, palette, pts, width, height) -> np.ndarray:
    S = np.ascontiguousarray(screen6, dtype=np.float32)
    pal = np.ascontiguousarray(palette, dtype=np.uint32)
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 4)
    hist = np.zeros((height, width, 4), dtype=np.uint32)
    clib().orc_bin_points(_ptr(S), _ptr(pal), _ptr(pts), pts.shape[0], _ptr(hist), width, height)
    return hist


def seed_walkers(params32, stride, blocks, tpb, width, height, seed, stream=0):
    n = blocks * tpb
    x, y, c = np.zeros(n), np.zeros(n), np.zeros(n)
    p = np.ascontiguousarray(params32, dtype=np.float32)
    clib().orc_seed(_ptr(p), stride, blocks, tpb, width, height, seed, stream, _ptr(x), _ptr(y), _ptr(c))
    return x, y, c


def trace(L: Layout, params32, stride, blocks, tpb, seed, pass0, npasses, x, y, c, gpu_trace=None, stream=0,
          last=None, gpu_xids=None):
    """per-pass walker states; xids per wave, or per walker for xaos flames
    (`last`: each walker's previous xform, 0xFFFFFFFF for none)"""
    prog = L.program()
    n = blocks * tpb
    x, y, c = (np.ascontiguousarray(v, dtype=np.float64).copy() for v in (x, y, c))
    tr = np.zeros((npasses, n, 4))
    xids = np.zeros((npasses, n if L.chaos else n // 64), dtype=np.uint32)
    p = np.ascontiguousarray(params32, dtype=np.float32)
    g = None if gpu_trace is None else np.ascontiguousarray(gpu_trace, dtype=np.float32)
    gx = None if gpu_xids is None else np.ascontiguousarray(gpu_xids, dtype=np.uint32)
    lst = None
    if L.chaos:
        lst = np.full(n, 0xFFFFFFFF, np.uint32) if last is None else np.ascontiguousarray(last, np.uint32).copy()
    rc = clib().orc_trace(_ptr(prog), L.n_xforms, int(L.has_final), L.chaos_base, _ptr(p), stride, blocks, tpb, seed,
                          stream, pass0, npasses, _ptr(x), _ptr(y), _ptr(c), None if lst is None else _ptr(lst),
                          _ptr(tr), _ptr(xids), None if g is None else _ptr(g), None if gx is None else _ptr(gx))
    assert rc == 0
    return tr, xids


def render(L: Layout, params32, stride, palettes, blocks, tpb, width, height, seed, fuse, passes, threads=0,
           stream=0):
    prog = L.program()
    hist = np.zeros((height, width, 4), dtype=np.uint32)
    p = np.ascontiguousarray(params32, dtype=np.float32)
    pal = np.ascontiguousarray(palettes, dtype=np.uint32)
    draws = clib().orc_render(_ptr(prog), L.n_xforms, int(L.has_final), L.chaos_base, _ptr(p), stride, _ptr(pal),
                              blocks, tpb, width, height, seed, stream, fuse, passes, _ptr(hist), threads)
    return hist, int(draws)


def render_f32(lib_path, L: Layout, params32, stride, palettes, blocks, tpb, width, height, seed, fuse, passes,
               threads=0, stream=0):
    """render() with the single-precision baseline build (build.build_f32)"""
    lib = C.CDLL(str(lib_path))
    P = C.c_void_p
    lib.orc_render.argtypes = [P, C.c_int, C.c_int, C.c_int, P, C.c_uint32, P, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, C.c_int]
    lib.orc_render.restype = C.c_uint64
    prog = L.program()
    hist = np.zeros((height, width, 4), dtype=np.uint32)
    p = np.ascontiguousarray(params32, dtype=np.float32)
    pal = np.ascontiguousarray(palettes, dtype=np.uint32)
    draws = lib.orc_render(_ptr(prog), L.n_xforms, int(L.has_final), L.chaos_base, _ptr(p), stride, _ptr(pal),
                           blocks, tpb, width, height, seed, stream, fuse, passes, _ptr(hist), threads)
    return hist, int(draws)


def prepare(flame: dict, config_dir, width: int, height: int, blocks: int, seed: int, stream: int = 0):
    """the oracle's own warmup: layout, 4 identical samples, Catmull-Rom,
    evaluation, precalc; returns (layout, float32 params [blocks][stride], palettes)"""
    vp = load_variation_params(config_dir)
    L = Layout(flame, vp)
    one = pack_sample(flame, L, width, height)
    samples = np.concatenate([one] * 4)
    ss = L.size_reals + 768
    fixed = fix_rotation(samples, ss, L.size_reals, L.affine_indices, 4)
    seg = catmull(fixed, ss)
    P, pals = eval_samples(seg, L, blocks)
    P = precalc(P, L, seed, stream)
    stride = param_stride(L.size_reals)
    P32 = np.zeros((blocks, stride), dtype=np.float32)
    P32[:, :L.size_reals] = P
    return L, P32, pals


def density_hdr(img: np.ndarray, bins97) -> np.ndarray:
    """CDF remap of a tonemapped RGBA image (assets/kernels/tonemap.cu:159-233):
    v *= (cdf(alpha * 96 + 1) / 0.5)^2 in HSV, alpha -> 1; the upper bin is
    clamped to the last one (the reference reads one past it at alpha = 1)"""
    img = np.asarray(img, dtype=np.float64)
    b = np.asarray(bins97, dtype=np.float64)
    cdf = (np.cumsum(b) / b.sum() if b.sum() > 0 else np.zeros(97)).astype(np.float32).astype(np.float64)
    power = img[..., 3] * 96.0 + 1.0
    lo = np.minimum(np.floor(power).astype(np.int64), 96)
    hi = np.minimum(lo + 1, 96)
    frac = power - np.floor(power)
    v = cdf[lo] + frac * (cdf[hi] - cdf[lo])
    r, g, bb = img[..., 0], img[..., 1], img[..., 2]
    mx = np.maximum(r, np.maximum(g, bb))
    mn = np.minimum(r, np.minimum(g, bb))
    d = mx - mn
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.where(mx == r, np.mod((g - bb) / d, 6.0), np.where(mx == g, (bb - r) / d + 2.0, (r - g) / d + 4.0))
        s = np.where(mx > 0, d / mx, 0.0)
    h = np.where(d > 0, h, 0.0)
    h = np.where(h < 0, h + 6.0, h)
    val = mx * (v / 0.5) ** 2
    i = np.floor(h).astype(np.int64) % 6
    f = h - np.floor(h)
    p, q, t = val * (1 - s), val * (1 - s * f), val * (1 - s * (1 - f))
    out = np.zeros(img.shape[:-1] + (4,))
    choices = [(val, t, p), (q, val, p), (p, val, t), (p, q, val), (t, p, val), (val, p, q)]
    for k, (cr, cg, cb) in enumerate(choices):
        m = i == k
        out[..., 0] = np.where(m, cr, out[..., 0])
        out[..., 1] = np.where(m, cg, out[..., 1])
        out[..., 2] = np.where(m, cb, out[..., 2])
    out[..., 3] = 1.0
    return out


def convert_image(img: np.ndarray, fmt: int, flip: bool = False, swap: bool = False) -> np.ndarray:
    """convert_to (assets/kernels/convert.cu:72-98) of a half3 / half4 image:
    element by element to float / half / 8-bit, where 8-bit is
    min(v * 255, 255) truncated toward zero (convert.cu:46-48; negative v,
    undefined there, give 0); a 3-channel input gets alpha 1; `swap`
    exchanges R and B; `flip` writes row y to row h-1-y. fmt: 0 float4,
    1 half4, 2 float3, 3 half3, 4 RGBA8, 5 RGB8."""
    h = np.asarray(img, dtype=np.float16)
    if h.shape[-1] == 3:
        h = np.concatenate([h, np.ones(h.shape[:-1] + (1,), np.float16)], axis=-1)
    if swap:
        h = h[..., [2, 1, 0, 3]]
    if flip:
        h = h[::-1]
    ch = 4 if fmt in (0, 1, 4) else 3
    h = h[..., :ch]
    if fmt in (0, 2):
        return h.astype(np.float32)
    if fmt in (1, 3):
        return h.copy()
    v = np.minimum(h.astype(np.float64) * 255.0, 255.0)
    return np.trunc(np.maximum(v, 0.0)).astype(np.uint8)


def jpeg_encode(img: np.ndarray, quality: int) -> bytes:
    """baseline JFIF of a uint8 h x w x 3|4 image (jpeg_oracle.c)"""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w, c = img.shape
    cap = 1 << 16
    while True:
        out = np.zeros(cap + 2 * 1024 * ((w + 7) // 8) * ((h + 7) // 8), dtype=np.uint8)
        n = clib().orc_jpeg_encode(_ptr(img), w * c, c, w, h, quality, _ptr(out), out.size)
        if n >= 0:
            return out[:n].tobytes()
        cap = -n

"""Variation programs pinned three ways, on the CPU:

1. the golden vectors (tests/golden/variation_vectors.npz: librefrakt's 99
   programs of config/variations_new.yml evaluated by the oracle's
   interpreter) against closed forms derived independently of the programs
   (complex-analytic identities and textbook formulas), which pins the
   interpreter and the fixture;
2. this build's variation database, interpreted, against the vectors: every
   one of the 99 programs must reproduce librefrakt's results;
3. the C oracle's hand-written formulas (used by its renders and trajectory
   checks) against the vectors.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import flang_eval as fe
from oracle import oracle

ROOT = Path(__file__).resolve().parents[1]
FIXTURE = ROOT / "tests" / "golden" / "variation_vectors.npz"
CONFIG = ROOT / "refraktor_amd" / "config"
REF_CONFIG = Path("/root/reference/config")

# agreement between two float64 evaluations of the same formula:
# |a - b| <= TOL * max(1, |b|), NaN / inf in the same places
TOL = 1e-9


@pytest.fixture(scope="module")
def vec():
    z = np.load(FIXTURE)
    return {"names": [str(n) for n in z["names"]], "params": json.loads(str(z["params"])), "px": z["px"],
            "py": z["py"], "words": z["words"], "weights": z["weights"], "affines": z["affines"], "rx": z["rx"],
            "ry": z["ry"]}


@pytest.fixture(scope="module")
def mine():
    return fe.load_db((CONFIG / "variations_new.yml").read_text(), (CONFIG / "common.yml").read_text())


def _aff(vec, s) -> dict:
    return dict(zip("abcdef", vec["affines"][s]))


def _close(gx, gy, rx, ry, tol=TOL, mask=None) -> tuple[bool, float]:
    g = np.concatenate([gx, gy])
    r = np.concatenate([rx, ry])
    m = np.ones_like(r, dtype=bool) if mask is None else np.concatenate([mask, mask])
    fin = np.isfinite(r) & np.isfinite(g)
    same_class = (np.isnan(g) == np.isnan(r)) & (np.isinf(g) == np.isinf(r))
    if not same_class[m].all():
        return False, float("inf")
    k = fin & m
    if not k.any():
        return True, 0.0
    err = np.abs(g[k] - r[k]) / np.maximum(1.0, np.abs(r[k]))
    return bool(err.max() <= tol), float(err.max())


def test_fixture_covers_the_reference_database(vec, mine):
    assert len(vec["names"]) == 99
    assert sorted(mine[0]) == vec["names"]


@pytest.mark.skipif(not REF_CONFIG.exists(), reason="reference checkout not present")
def test_fixture_is_what_the_reference_programs_give(vec):
    """regenerating from librefrakt's YAML reproduces the committed vectors"""
    db, commons = fe.load_db((REF_CONFIG / "variations_new.yml").read_text(), (REF_CONFIG / "common.yml").read_text())
    for i, name in enumerate(vec["names"]):
        for s in range(3):
            rx, ry = fe.evaluate(db[name], commons, vec["px"], vec["py"], vec["weights"][s], vec["params"][name][s],
                                 _aff(vec, s), words=vec["words"].astype(np.uint64))
            ok, err = _close(rx, ry, vec["rx"][i, s], vec["ry"][i, s], tol=0.0)
            assert ok, (name, s, err)


# ------------------------------------------------------------ closed forms

def _u01(words, k):
    """the k-th uniform of each point: float32(u) * 2^-32 (random.h:16-22)"""
    return words[:, k].astype(np.float32).astype(np.float64) * 2.0 ** -32


def _cz(x, y):
    return x + 1j * y


def _closed_forms():
    """name -> f(z, x, y, w, prm, aff, words) giving the complex result"""
    def r_(z):
        return np.abs(z)

    def atn(x, y):      # librefrakt's common.atan: atan2(x, y)
        return np.arctan2(x, y)

    cf = {
        "linear": lambda z, x, y, w, p, a, u: w * z,
        "sinusoidal": lambda z, x, y, w, p, a, u: w * _cz(np.sin(x), np.sin(y)),
        "spherical": lambda z, x, y, w, p, a, u: w * z / np.abs(z) ** 2,
        # x' + i y' = z (sin r^2 + i cos r^2)
        "swirl": lambda z, x, y, w, p, a, u: w * z * (np.sin(np.abs(z) ** 2) + 1j * np.cos(np.abs(z) ** 2)),
        "horseshoe": lambda z, x, y, w, p, a, u: w * z * z / np.abs(z),
        "polar": lambda z, x, y, w, p, a, u: w * _cz(atn(x, y) / np.pi, np.abs(z) - 1.0),
        "handkerchief": lambda z, x, y, w, p, a, u: w * r_(z) * _cz(np.sin(atn(x, y) + r_(z)), np.cos(atn(x, y) - r_(z))),
        "heart": lambda z, x, y, w, p, a, u: w * r_(z) * _cz(np.sin(r_(z) * atn(x, y)), -np.cos(r_(z) * atn(x, y))),
        "disc": lambda z, x, y, w, p, a, u: w * atn(x, y) / np.pi * _cz(np.sin(np.pi * r_(z)), np.cos(np.pi * r_(z))),
        "spiral": lambda z, x, y, w, p, a, u: w / r_(z) * _cz(y / r_(z) + np.sin(r_(z)), x / r_(z) - np.cos(r_(z))),
        "hyperbolic": lambda z, x, y, w, p, a, u: w * _cz(x / r_(z) ** 2, y),
        "diamond": lambda z, x, y, w, p, a, u: w * _cz(x / r_(z) * np.cos(r_(z)), y / r_(z) * np.sin(r_(z))),
        # julia: sqrt of (y + i x) scaled, sign from bit 24 of the first word
        "julia": lambda z, x, y, w, p, a, u: (w * np.sqrt(r_(z)) * np.exp(0.5j * atn(x, y))
                                             * np.where((u[:, 0] >> 24) & 1, -1.0, 1.0)),
        "bubble": lambda z, x, y, w, p, a, u: w * z / (0.25 * np.abs(z) ** 2 + 1.0),
        "curl": lambda z, x, y, w, p, a, u: w * z / (1.0 + p["c1"] * z + p["c2"] * z * z),
        "fisheye": lambda z, x, y, w, p, a, u: 2.0 * w / (r_(z) + 1.0) * _cz(y, x),
        "eyefish": lambda z, x, y, w, p, a, u: 2.0 * w / (r_(z) + 1.0) * z,
        "cylinder": lambda z, x, y, w, p, a, u: w * _cz(np.sin(x), y),
        "exponential": lambda z, x, y, w, p, a, u: w * np.exp(x - 1.0) * np.exp(1j * np.pi * y),
        "cosine": lambda z, x, y, w, p, a, u: w * np.cos(np.pi * x + 1j * y),
        "exp": lambda z, x, y, w, p, a, u: w * np.exp(z),
        "log": lambda z, x, y, w, p, a, u: w * np.log(z),
        "sin": lambda z, x, y, w, p, a, u: w * np.sin(z),
        # librefrakt's cos is the conjugate of the complex cosine
        "cos": lambda z, x, y, w, p, a, u: w * np.conj(np.cos(z)),
        "cosh": lambda z, x, y, w, p, a, u: w * np.cosh(z),
        # librefrakt's sinh takes p.y for both parts: sinh(y + i y)
        "sinh": lambda z, x, y, w, p, a, u: w * np.sinh(_cz(y, y)),
        "tan": lambda z, x, y, w, p, a, u: w * np.tan(z),
        "tanh": lambda z, x, y, w, p, a, u: w * np.tanh(z),
        "sec": lambda z, x, y, w, p, a, u: w / np.cos(z),
        "csc": lambda z, x, y, w, p, a, u: w / np.sin(z),
        "cot": lambda z, x, y, w, p, a, u: w / np.tan(z),
        "sech": lambda z, x, y, w, p, a, u: w / np.cosh(z),
        "csch": lambda z, x, y, w, p, a, u: w / np.sinh(z),
        "coth": lambda z, x, y, w, p, a, u: w / np.tanh(z),
        "polar2": lambda z, x, y, w, p, a, u: w / np.pi * _cz(atn(x, y), np.log(np.abs(z))),
        "power": lambda z, x, y, w, p, a, u: w * r_(z) ** (x / r_(z)) * _cz(y / r_(z), x / r_(z)),
        "blur": lambda z, x, y, w, p, a, u: w * _u01(u, 0) * np.exp(2j * np.pi * _u01(u, 1)),
        "square": lambda z, x, y, w, p, a, u: w * _cz(_u01(u, 0) - 0.5, _u01(u, 1) - 0.5),
        "noise": lambda z, x, y, w, p, a, u: (w * _u01(u, 0) * _cz(x * np.cos(2 * np.pi * _u01(u, 1)),
                                                                  y * np.sin(2 * np.pi * _u01(u, 1)))),
        "gaussian_blur": lambda z, x, y, w, p, a, u: (w * (_u01(u, 1) + _u01(u, 2) + _u01(u, 3) + _u01(u, 4) - 2.0)
                                                      * np.exp(2j * np.pi * _u01(u, 0))),
        "cross": lambda z, x, y, w, p, a, u: w * z / np.abs(x * x - y * y),
        "foci": lambda z, x, y, w, p, a, u: (w / (np.cosh(x) - np.cos(y))
                                             * _cz(np.sinh(x), np.sin(y))),
        # julian: a branch of z^(1/power) with modulus |z|^(dist/power)
        "julian": lambda z, x, y, w, p, a, u: (w * np.abs(z) ** (p["dist"] / p["power"])
                                               * np.exp(1j * (np.angle(z) + 2 * np.pi * np.trunc(abs(p["power"]) * _u01(u, 0)))
                                                        / p["power"])),
        "waves": lambda z, x, y, w, p, a, u: w * _cz(x + a["b"] * np.sin(y / a["c"] ** 2),
                                                      y + a["e"] * np.sin(x / a["f"] ** 2)),
        "popcorn": lambda z, x, y, w, p, a, u: w * _cz(x + a["c"] * np.sin(np.tan(3 * y)),
                                                        y + a["f"] * np.sin(np.tan(3 * x))),
        "pdj": lambda z, x, y, w, p, a, u: w * _cz(np.sin(p["a"] * y) - np.cos(p["b"] * x),
                                                    np.sin(p["c"] * x) - np.cos(p["d"] * y)),
        "bent": lambda z, x, y, w, p, a, u: w * _cz(np.where(x < 0, 2 * x, x), np.where(y < 0, y / 2, y)),
        "splits": lambda z, x, y, w, p, a, u: w * _cz(x + np.where(x >= 0, p["x"], -p["x"]),
                                                       y + np.where(y >= 0, p["y"], -p["y"])),
        "loonie": lambda z, x, y, w, p, a, u: np.where(np.abs(z) < abs(w), w * np.sqrt(w * w / np.abs(z) ** 2 - 1.0),
                                                      w) * z,
        # escher: z^(vc + i vd) with vc = (1 + cos beta)/2, vd = sin(beta)/2
        "escher": lambda z, x, y, w, p, a, u: w * np.exp((0.5 * (1 + np.cos(p["beta"])) + 0.5j * np.sin(p["beta"]))
                                                          * np.log(z)),
        # cpow: z^((r + i i)/power) on the branch chosen by the first draw
        "cpow": lambda z, x, y, w, p, a, u: (w * np.exp((p["r"] + 1j * p["i"]) / p["power"] * np.log(z))
                                             * np.exp(2j * np.pi / p["power"] * np.floor(p["power"] * _u01(u, 0)))),
    }
    return cf


CLOSED = _closed_forms()


@pytest.mark.parametrize("name", sorted(CLOSED))
def test_reference_vectors_match_closed_forms(vec, name):
    i = vec["names"].index(name)
    px, py = vec["px"], vec["py"]
    z = _cz(px, py)
    words = vec["words"].astype(np.uint64)
    for s in range(3):
        prm = vec["params"][name][s]
        with np.errstate(all="ignore"):
            out = CLOSED[name](z, px, py, float(vec["weights"][s]), prm, _aff(vec, s), words)
        rx, ry = vec["rx"][i, s], vec["ry"][i, s]
        # closed forms are compared where the program's result is finite and
        # away from the branch cuts and poles the two forms treat differently
        with np.errstate(all="ignore"):
            m = np.isfinite(rx) & np.isfinite(ry) & (np.abs(z) > 1e-5) & (np.abs(out) < 1e6)
        if name in ("julian", "cpow", "escher", "log"):
            m &= np.abs(py) > 1e-9  # principal-argument cut on the negative real axis
        if name in ("tan", "cot", "sec", "csc", "tanh", "coth", "sech", "csch", "cross", "foci"):
            m &= np.abs(out) < 1e3  # near poles
        assert m.mean() > 0.8, (name, s, m.mean())
        err = np.abs(out[m] - (rx[m] + 1j * ry[m])) / np.maximum(1.0, np.abs(out[m]))
        assert err.max() < 1e-9, (name, s, float(err.max()), int(np.argmax(err)))


def test_closed_forms_catch_planted_mistakes(vec):
    """a dropped term, a wrong sign or swapped outputs fail the closed form"""
    i = vec["names"].index("curl")
    px, py = vec["px"], vec["py"]
    z = _cz(px, py)
    prm = vec["params"]["curl"][1]
    w = float(vec["weights"][1])
    good = CLOSED["curl"](z, px, py, w, prm, None, None)
    ref = vec["rx"][i, 1] + 1j * vec["ry"][i, 1]
    assert np.max(np.abs(good - ref) / np.maximum(1, np.abs(good))) < 1e-9
    for bad in (w * z / (1.0 + prm["c1"] * z),                       # dropped c2 term
                w * z / (1.0 - prm["c1"] * z + prm["c2"] * z * z),    # wrong sign
                1j * np.conj(good)):                                  # x and y swapped
        assert np.max(np.abs(bad - ref) / np.maximum(1, np.abs(good))) > 1e-3


# ------------------------------------------------------ this build's programs

def test_database_programs_match_reference_vectors(vec, mine):
    db, commons = mine
    failures = []
    for i, name in enumerate(vec["names"]):
        for s in range(3):
            rx, ry = fe.evaluate(db[name], commons, vec["px"], vec["py"], vec["weights"][s], vec["params"][name][s],
                                 _aff(vec, s), words=vec["words"].astype(np.uint64))
            ok, err = _close(rx, ry, vec["rx"][i, s], vec["ry"][i, s])
            if not ok:
                failures.append((name, s, err))
    assert not failures, failures


def test_database_random_draw_order(vec, mine):
    """variations that draw consume the words in librefrakt's order: a
    shifted word stream changes their results"""
    db, commons = mine
    shifted = np.roll(vec["words"].astype(np.uint64), 1, axis=1)
    for name in vec["names"]:
        v = db[name]
        if not fe.uses_random(v):
            continue
        i = vec["names"].index(name)
        same = []
        for s in range(3):
            rx, ry = fe.evaluate(v, commons, vec["px"], vec["py"], vec["weights"][s], vec["params"][name][s],
                                 _aff(vec, s), words=shifted)
            same.append(_close(rx, ry, vec["rx"][i, s], vec["ry"][i, s])[0])
        assert not all(same), name


# ---------------------------------------------------------------- C oracle

@pytest.mark.parametrize("name", oracle.VARIATIONS)
def test_c_oracle_formulas_match_reference_vectors(vec, mine, name):
    db, _ = mine
    i = vec["names"].index(name)
    for s in range(3):
        aff = _aff(vec, s)
        full = fe.precalc(db[name], vec["params"][name][s], aff)
        rx, ry = oracle.eval_variation(name, float(vec["weights"][s]), full, aff, vec["px"], vec["py"], vec["words"])
        ok, err = _close(rx, ry, vec["rx"][i, s], vec["ry"][i, s])
        assert ok, (name, s, err)

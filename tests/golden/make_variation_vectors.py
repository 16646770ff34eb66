"""Writes tests/golden/variation_vectors.npz: input/output vectors of every
variation program of librefrakt's config/variations_new.yml and
config/common.yml, evaluated by the oracle's interpreter of the variation
language (oracle/flang_eval.py) in float64.

Run in a container that holds the reference checkout:
    python tests/golden/make_variation_vectors.py [/root/reference]

The fixture is data only: the points, the random words each point may
consume, the parameter sets, and the results. Inputs:
  * 512 points shared by every variation: 432 uniform in [-2.5, 2.5]^2, 48
    within 0.05 of the origin, and 32 special points (origin, axes,
    diagonals, a few large ones);
  * 16 random 32-bit words per point (numpy PCG64, seed 20240611);
  * three (weight, vlink affine, parameter) sets per variation: weight 0.8,
    the YAML defaults with zero defaults replaced by 0.7; weight 1.3 and
    weight -0.6 with every user parameter drawn from +-U[0.2, 2.0].
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import flang_eval as fe  # noqa: E402

OUT = HERE / "variation_vectors.npz"
SEED = 20240611
N_WORDS = 16
AFFINES = [(0.9, 0.3, 0.4, -0.2, 1.1, -0.35), (1.2, -0.5, -0.6, 0.4, 0.7, 0.25),
           (-0.8, 0.2, 0.15, 0.6, -1.0, 0.5)]   # (a, b, c, d, e, f)
WEIGHTS = [0.8, 1.3, -0.6]


def points(rng: np.random.Generator) -> tuple[np.ndarray, np.ndarray]:
    px = list(rng.uniform(-2.5, 2.5, 432))
    py = list(rng.uniform(-2.5, 2.5, 432))
    r = rng.uniform(0.0, 0.05, 48)
    a = rng.uniform(-np.pi, np.pi, 48)
    px += list(r * np.cos(a))
    py += list(r * np.sin(a))
    special = [(0.0, 0.0), (1.0, 0.0), (-1.0, 0.0), (0.0, 1.0), (0.0, -1.0), (0.5, 0.5), (-0.5, 0.5),
               (0.5, -0.5), (-0.5, -0.5), (2.0, 0.0), (0.0, -2.0), (1e-3, 0.0), (0.0, 1e-3), (3.0, -2.0),
               (-4.0, 1.5), (6.0, 5.0), (0.25, 0.0), (0.0, 0.25), (-0.25, 1.75), (1.5, 1.5), (-1.5, -1.5),
               (0.7, -0.1), (-0.1, 0.7), (2.2, 2.4), (1e-6, 1e-6), (-1e-6, 2e-6), (0.3, 0.4), (0.6, 0.8),
               (1.0, 1.0), (-1.0, -1.0), (1.0, -1.0), (-1.0, 1.0)]
    px += [s[0] for s in special]
    py += [s[1] for s in special]
    return np.array(px), np.array(py)


def param_sets(v: fe.Variation, rng: np.random.Generator) -> list[dict]:
    user = v.user_params
    sets = [{k: (v.params[k]["default"] if v.params[k]["default"] != 0.0 else 0.7) for k in user}]
    for _ in range(2):
        sets.append({k: float(np.round(rng.choice([-1.0, 1.0]) * rng.uniform(0.2, 2.0), 4)) for k in user})
    return sets


def main(ref_root: str = "/root/reference") -> None:
    cfg = Path(ref_root) / "config"
    db, commons = fe.load_db((cfg / "variations_new.yml").read_text(), (cfg / "common.yml").read_text())
    rng = np.random.default_rng(SEED)
    px, py = points(rng)
    words = rng.integers(0, 1 << 32, size=(px.size, N_WORDS), dtype=np.uint64).astype(np.uint32)
    names = sorted(db)
    rx = np.zeros((len(names), 3, px.size))
    ry = np.zeros_like(rx)
    params = {}
    for i, name in enumerate(names):
        v = db[name]
        params[name] = param_sets(v, rng)
        for s in range(3):
            aff = dict(zip("abcdef", AFFINES[s]))
            rx[i, s], ry[i, s] = fe.evaluate(v, commons, px, py, WEIGHTS[s], params[name][s], aff,
                                             words=words.astype(np.uint64))
    np.savez_compressed(OUT, names=np.array(names), params=np.array(json.dumps(params)), px=px, py=py,
                        words=words, weights=np.array(WEIGHTS), affines=np.array(AFFINES), rx=rx, ry=ry)
    print(f"wrote {OUT} ({len(names)} variations x 3 sets x {px.size} points)")


if __name__ == "__main__":
    main(*sys.argv[1:2])

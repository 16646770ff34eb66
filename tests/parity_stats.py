"""Converged-image comparison shared by the CPU and GPU parity tests.

The reference's histogram is built with racy float adds (refactor.cu:366-381),
so renders can only agree statistically. Two renders are compared per pixel:
over the pixels where the reference render holds at least `min_draws`
full-opacity draws, the mean of sum_c |a_c - b_c| / sum_c b_c (c = R, G, B, W
fixed-point channels). The GPU render and the oracle render it is compared
with use independent streams (different seeds), and the statistic is divided
by the same statistic between two oracle renders of other independent seeds
at the same budget: a correct renderer gives a ratio near 1 (0.997 and 1.012
on the 640x360 C1 flames of seeds 3 and 6 at 2^26 iterations), and a 0.5 %
change of one affine coefficient, a 2 % change of a variation weight or a 3 %
change of an xform colour, each in the lightest xform, gives 1.3 to 6.2.
"""
from __future__ import annotations

import numpy as np

TAU = 1.15          # accepted ratio to the oracle's own seed-to-seed statistic
MIN_DRAWS = 256


def pixel_rel_l1(a: np.ndarray, ref: np.ndarray, min_draws: int = MIN_DRAWS) -> tuple[float, int]:
    """(mean per-pixel relative L1 of a against ref, pixels compared)"""
    a = a.astype(np.float64).reshape(-1, 4)
    b = ref.astype(np.float64).reshape(-1, 4)
    m = b[:, 3] >= 255.0 * min_draws
    d = np.abs(a[m] - b[m]).sum(axis=1) / b[m].sum(axis=1)
    return float(d.mean()) if m.any() else float("nan"), int(m.sum())


def converged_ratio(test: np.ndarray, ref: np.ndarray, ref_other: np.ndarray,
                    min_draws: int = MIN_DRAWS) -> tuple[float, float, float, int]:
    """(ratio, statistic of test vs ref, statistic of ref_other vs ref,
    pixels): test and ref_other must come from streams independent of ref's"""
    s_test, n = pixel_rel_l1(test, ref, min_draws)
    s_seed, _ = pixel_rel_l1(ref_other, ref, min_draws)
    return s_test / s_seed, s_test, s_seed, n

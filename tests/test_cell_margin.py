"""The cell grid's edge bound (CPU, no GPU).

The cell search needs its cell edge to bound the widest AABB extent of the
*shifted* coordinates, computed exactly from the stored corners (the search's
overlap test runs on those).  The grid is planned before the shift, from W,
the widest extent the bounds pass computes (float differences of the
unshifted corners, cast to float first), and C, the largest coordinate
magnitude of the scene, shifted or not:

    exact shifted extent <= W * (1 + 2^-23) + 2^-22 * C

and the edge is (1 + 1e-3) * (W + 2^-20 * C).  These tests evaluate the
same roundings with numpy (float32 scenes: the corners, their shift by the
scene origin, the float difference; float64 scenes: the double shift and the
cast of the corners to float) over random and adversarial scenes, and check
the bound and the edge against the exact extent.  A margin without the C
term fails the far-from-origin cases.
"""
import numpy as np
import pytest


def _bound_terms(lo, hi, dtype):
    """W (float32, as the bounds pass computes it), C, and the exact widest
    shifted extent for a scene with corners lo, hi (arrays of dtype)."""
    lo = lo.astype(dtype)
    hi = hi.astype(dtype)
    # bounds pass: float corners, float differences (NaN-free inputs here)
    lo_f, hi_f = lo.astype(np.float32), hi.astype(np.float32)
    W = float(np.max(hi_f - lo_f))
    # scene bounds over both corners as floats; origin = |min| (Broadphase.cu:377-381)
    bmin = np.minimum(lo_f.min(axis=0), hi_f.min(axis=0)).astype(np.float64)
    bmax = np.maximum(lo_f.max(axis=0), hi_f.max(axis=0)).astype(np.float64)
    origin = np.abs(bmin).astype(dtype)
    # shift in the scene's real type (one addition per coordinate)
    slo = (lo + origin).astype(dtype)
    shi = (hi + origin).astype(dtype)
    exact = float(np.max(shi.astype(np.float64) - slo.astype(np.float64)))
    ext = np.maximum(bmax + origin.astype(np.float64), 0.0)
    C = float(max(np.abs(bmin).max(), np.abs(bmax).max(), ext.max()))
    return W, C, exact


def _scene(rng, n, center, spread, wmin, wmax):
    c = center + rng.uniform(-spread, spread, (n, 3))
    w = rng.uniform(wmin, wmax, (n, 3))
    return c - w / 2, c + w / 2


CASES = [
    # (centre, spread, min width, max width)
    (0.0, 1.0, 1e-3, 2e-2),          # lattice-like around the origin
    (1.0e4, 5.0, 0.05, 0.15),        # far from the origin, positive
    (-1.0e4, 5.0, 0.05, 0.15),       # far from the origin, negative
    (3.0e5, 1.0, 1e-4, 1e-3),        # tiny boxes far out: rounding dominates W
    (-7.5e6, 10.0, 1e-2, 1.0),
    (0.0, 1.0e3, 0.0, 0.0),          # zero-width boxes
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", CASES)
def test_shifted_extent_within_bound(dtype, case):
    rng = np.random.default_rng(2 * CASES.index(case) + (dtype == np.float64))
    for _ in range(20):
        lo, hi = _scene(rng, 4000, *case)
        W, C, exact = _bound_terms(lo, hi, dtype)
        bound = W * (1 + 2.0 ** -23) + 2.0 ** -22 * C
        assert exact <= bound, (exact, bound, W, C)
        edge = (1 + 1e-3) * (W + 2.0 ** -20 * C + 8.0 * np.finfo(np.float32).smallest_subnormal)
        # the search's argument: the exact extent stays below edge / (1 + 1e-3) * (1 + 1e-6)
        assert exact <= edge / (1 + 1e-3) * (1 + 1e-6)


def test_margin_needs_the_coordinate_term():
    """Without the 2^-20 C term the edge can fall below the shifted extent:
    tiny float32 boxes far from the origin grow when shifted."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(50):
        lo, hi = _scene(rng, 4000, 3.0e5, 1.0, 1e-4, 1e-3)
        W, C, exact = _bound_terms(lo, hi, np.float32)
        if W > 0:
            worst = max(worst, exact / ((1 + 1e-3) * W))
        elif exact > 0:
            worst = np.inf
    assert worst > 1.0

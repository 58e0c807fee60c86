"""Pins the oracle's ray/sphere primitives (restating sim3d_kernel.cu:77-160)
against closed forms, the branch rules of the tolerance band, and brute-force
root finding."""
import numpy as np
import pytest

TOL = float(np.float32(1e-4))  # sim3d_kernel.cu:34, `const float tol = 1e-4`


def test_inside_axis_aligned(oracle_lib):
    # spin at the centre of a sphere of radius 1, jump of length 2 along +x:
    # the exit shell r - tol is reached at fraction (1 - tol) / 2
    f = oracle_lib.step_inside(0.0, 0.0, 0.0, 1.0, 2.0, 0.0, 0.0)
    assert f == pytest.approx((1 - TOL) / 2, rel=1e-14)


def test_inside_offset_point(oracle_lib):
    # d = (0.5, 0, 0), j = (-1, 0, 0): exit at x = -(1 - tol) -> f = 1.5 - tol
    f = oracle_lib.step_inside(0.5, 0.0, 0.0, 1.0, -1.0, 0.0, 0.0)
    assert f == pytest.approx(1.5 - TOL, rel=1e-14)


def test_inside_returns_large_when_outside(oracle_lib):
    assert oracle_lib.step_inside(2.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0) == 1e6


def test_inside_band_pointing_out_returns_zero(oracle_lib):
    # spin inside r but outside r - tol, moving outward: both roots <= 0 -> 0
    r = 1.0
    x = r - TOL / 2
    assert oracle_lib.step_inside(x, 0.0, 0.0, r, 0.3, 0.0, 0.0) == 0.0


def test_inside_band_pointing_in_crosses(oracle_lib):
    # same spin moving inward: the largest positive root (far side) is taken
    r = 1.0
    x = r - TOL / 2
    f = oracle_lib.step_inside(x, 0.0, 0.0, r, -1.0, 0.0, 0.0)
    assert f == pytest.approx(x + (r - TOL), rel=1e-12)


def test_outside_head_on(oracle_lib):
    # d = (3, 0, 0), j = (-2, 0, 0): entry shell r + tol at x = 1 + tol
    f = oracle_lib.step_outside(3.0, 0.0, 0.0, 1.0, -2.0, 0.0, 0.0)
    assert f == pytest.approx((2 - TOL) / 2, rel=1e-14)


def test_outside_miss_and_away(oracle_lib):
    assert oracle_lib.step_outside(3.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0) == 1e6   # disc < 0
    assert oracle_lib.step_outside(3.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0) == 1e6   # both roots < 0


def test_outside_near_tangent_rays(oracle_lib):
    # ray along +y at x just inside the r + tol shell grazes it: entry at
    # f = 2 - sqrt(R^2 - x^2); just outside the shell it misses
    R = 1.0 + TOL
    x = R - 1e-6
    f = oracle_lib.step_outside(x, -2.0, 0.0, 1.0, 0.0, 1.0, 0.0)
    assert f == pytest.approx(2.0 - np.sqrt(R * R - x * x), rel=1e-6)
    assert oracle_lib.step_outside(R + 1e-6, -2.0, 0.0, 1.0, 0.0, 1.0, 0.0) == 1e6


def test_outside_band_branches(oracle_lib):
    # spin between r and r + tol: roots of opposite sign.  Moving towards the
    # sphere (b < 0) stops at once (0); moving away (b > 0) is free (1e6).
    r = 1.0
    x = r + TOL / 2
    assert oracle_lib.step_outside(x, 0.0, 0.0, r, -0.5, 0.0, 0.0) == 0.0
    assert oracle_lib.step_outside(x, 0.0, 0.0, r, 0.5, 0.0, 0.0) == 1e6


def _bisect_first_crossing(d, j, R, fmax):
    """Smallest f in [0, fmax] where |d + f j| crosses R, by dense sampling
    plus bisection."""
    fs = np.linspace(0, fmax, 400001)
    pts = d[None, :] + fs[:, None] * j[None, :]
    g = np.linalg.norm(pts, axis=1) - R
    idx = np.nonzero(np.sign(g[1:]) != np.sign(g[:-1]))[0]
    if idx.size == 0:
        return None
    lo, hi = fs[idx[0]], fs[idx[0] + 1]
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        gm = np.linalg.norm(d + mid * j) - R
        if np.sign(gm) == np.sign(np.linalg.norm(d + lo * j) - R):
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_random_rays_against_bisection(oracle_lib):
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        r = rng.uniform(0.3, 2.0)
        d = rng.normal(size=3) * r
        j = rng.normal(size=3) * rng.uniform(0.2, 3.0)
        inside = np.linalg.norm(d) < r
        if inside:
            if np.linalg.norm(d) > r - TOL:
                continue
            f = oracle_lib.step_inside(*d, r, *j)
            ref = _bisect_first_crossing(d, j, r - TOL, 3 * r / np.linalg.norm(j) + 1)
            assert ref is not None and f == pytest.approx(ref, rel=1e-6, abs=1e-9)
        else:
            if np.linalg.norm(d) < r + TOL:
                continue
            f = oracle_lib.step_outside(*d, r, *j)
            ref = _bisect_first_crossing(d, j, r + TOL, 4.0)
            if ref is None:
                assert f == 1e6 or f > 4
            else:
                assert f == pytest.approx(ref, rel=1e-6, abs=1e-9)
        checked += 1
    assert checked > 200

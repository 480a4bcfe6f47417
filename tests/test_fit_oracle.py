"""Pins the fit half of the CPU oracle (oracle/fit_oracle.c), which restates
the reference's src/fit.rs.

* The reference's own fit tests (src/fit.rs:566-704), run on the oracle:
  spline on linear data, the initial simplex inside custom bounds, the
  sampled grid minimum (1.25, 0.04) with chi2 < 1e-24, custom c2 bounds
  (1.25, 1.2), and the validation errors.
* Closed forms the mathematics fixes: a Nelder-Mead problem whose minimum
  is known exactly (total = [1, c1, 10 c2] against data generated at an
  interior point, and against a target outside the box, where the fit must
  stop on the bound).
* Independent implementations: the natural cubic spline against
  scipy.interpolate.CubicSpline(bc_type="natural"); the analytic scale and
  chi2/N against a least-squares solve (numpy.linalg.lstsq) on scipy's spline
  values; the grid fit against a brute-force numpy scan of the same grid.
"""
from __future__ import annotations

import numpy as np
import pytest

Q3 = np.array([0.0, 1.0, 2.0])


def theory3(c1, c2):
    """The reference tests' closure: total = [1, c1, 10 c2] at q = [0, 1, 2]."""
    return np.array([1.0, c1, 10.0 * c2])


def test_spline_matches_linear_on_linear_data(oracle):
    # src/fit.rs:566-575, exact
    got = oracle.spline_eval([0.0, 1.0, 2.0], [10.0, 20.0, 30.0], [-1.0, 3.0, 0.5, 1.5])
    assert got.tolist() == [10.0, 30.0, 15.0, 25.0]


def test_spline_matches_scipy_natural_spline(oracle):
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(3)
    for n in (3, 4, 7, 40, 301):
        xs = np.sort(rng.uniform(0.0, 0.6, n))
        xs = np.unique(xs)
        ys = rng.normal(size=len(xs)) * 1e3 + 5e3
        x = np.concatenate([rng.uniform(xs[0], xs[-1], 200), xs])
        want = CubicSpline(xs, ys, bc_type="natural")(x)
        got = oracle.spline_eval(xs, ys, x)
        assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(ys)), n
        # outside the table the reference returns the end values
        out = oracle.spline_eval(xs, ys, [xs[0] - 1.0, xs[-1] + 1.0])
        assert out.tolist() == [ys[0], ys[-1]]


def test_spline_two_points_is_linear(oracle):
    # n < 3: all second derivatives stay zero (src/fit.rs:149-156)
    got = oracle.spline_eval([1.0, 3.0], [2.0, 6.0], [1.5, 2.0, 2.5])
    assert np.allclose(got, [3.0, 4.0, 5.0], rtol=0, atol=1e-15)


def test_scale_and_chi2_against_least_squares(oracle):
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(11)
    tq = np.linspace(0.0, 0.5, 60)
    ttot = 1e4 * np.exp(-30.0 * tq * tq) + 50.0
    q = np.sort(rng.uniform(0.01, 0.49, 45))
    ith = CubicSpline(tq, ttot, bc_type="natural")(q)
    inten = 2.5e-3 * ith * (1.0 + 0.05 * rng.normal(size=len(q)))
    sigma = 0.02 * inten + 1e-3
    A = (ith / sigma)[:, None]
    b = inten / sigma
    sol, res, _, _ = np.linalg.lstsq(A, b, rcond=None)
    scale = oracle.compute_scale(tq, ttot, q, inten, sigma)
    chi2 = oracle.chi2_per_point(tq, ttot, q, inten, sigma)
    assert abs(scale - sol[0]) <= 1e-12 * abs(sol[0])
    assert abs(chi2 - res[0] / len(q)) <= 1e-9 * res[0] / len(q)
    # a theory that is the data times a constant fits exactly
    assert oracle.chi2_per_point(q, 4.0 * inten, q, inten, sigma) < 1e-28
    assert abs(oracle.compute_scale(q, 4.0 * inten, q, inten, sigma) - 0.25) < 1e-15


def test_initial_simplex_respects_custom_bounds(oracle):
    # src/fit.rs:587-597
    s = oracle.initial_simplex((2.0, 3.0), (0.2, 0.4))
    for c1, c2 in s:
        assert 2.0 <= c1 <= 3.0 and 0.2 <= c2 <= 0.4
    assert s[0] != s[1] and s[0] != s[2]
    # default bounds: (1, 0.03), (1 + 0.08, 0.03), (1, 0.03 - 0.015)
    assert oracle.initial_simplex((0.5, 2.0), (0.0, 0.1)) == [(1.0, 0.03), (1.08, 0.03),
                                                               (1.0, 0.015)]
    # offsets flip to the other side at a bound and shrink to half the box
    s = oracle.initial_simplex((0.5, 1.02), (0.03, 0.04))
    assert s == [(1.0, 0.03), (1.0 - 0.08, 0.03), (1.0, 0.03 + (0.04 - 0.03) / 2.0)]


def test_nelder_mead_rejects_invalid_bounds(oracle):
    # src/fit.rs:599-618
    with pytest.raises(ValueError, match="lower bound"):
        oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 0.4], c1_bounds=(2.0, 0.5))
    with pytest.raises(ValueError, match="c2 bounds must be finite"):
        oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 0.4], c2_bounds=(0.0, np.inf))


def test_grid_fit_finds_sampled_minimum(oracle):
    # src/fit.rs:620-640
    c1, c2, chi2, scale = oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 0.4], method="grid",
                                     c1_points=7, c2_points=6)
    assert abs(c1 - 1.25) < 1e-12 and abs(c2 - 0.04) < 1e-12
    assert chi2 < 1e-24
    assert abs(scale - 1.0) < 1e-12


def test_grid_fit_rejects_degenerate_axes(oracle):
    # src/fit.rs:642-660
    with pytest.raises(ValueError, match="at least 2 points"):
        oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 0.4], method="grid", c1_points=1, c2_points=6)


def test_grid_fit_accepts_custom_c2_bounds(oracle):
    # src/fit.rs:662-683
    c1, c2, chi2, _ = oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 12.0], method="grid",
                                 c1_points=7, c2_points=6, c1_bounds=(0.5, 2.0),
                                 c2_bounds=(0.0, 2.0))
    assert abs(c1 - 1.25) < 1e-12 and abs(c2 - 1.2) < 1e-12
    assert chi2 < 1e-24


def test_grid_fit_rejects_invalid_bounds(oracle):
    # src/fit.rs:685-703
    with pytest.raises(ValueError, match="lower bound"):
        oracle.fit(theory3, Q3, Q3, [1.0, 1.25, 12.0], method="grid", c1_points=7,
                   c2_points=6, c1_bounds=(0.5, 2.0), c2_bounds=(2.0, 0.0))


@pytest.mark.parametrize("target", [(1.25, 0.04), (0.71, 0.083), (1.93, 0.002)])
def test_nelder_mead_recovers_interior_minimum(oracle, target):
    """chi2/N = 0 exactly at the generating point, which lies inside the box;
    the simplex must contract onto it.  The reference stops once the spread
    of chi2/N over the simplex is below 1e-10 (src/fit.rs:263-326), so that
    is how close the result gets: here chi2/N is |delta_perp|^2 / 3 (the
    part of the theory error not absorbed by the scale), so chi2/N < 1e-10
    bounds the parameter error to ~1e-5 in c1 and ~1e-6 in c2 (x10 in the
    theory)."""
    data = theory3(*target) * 1.0
    c1, c2, chi2, scale = oracle.fit(theory3, Q3, Q3, data)
    assert chi2 < 1e-10
    assert abs(c1 - target[0]) < 3e-5 and abs(c2 - target[1]) < 3e-6
    assert abs(scale - 1.0) < 3e-5


def test_nelder_mead_start_at_minimum(oracle):
    """The first vertex of the initial simplex, (1, 0.03), is the minimum:
    it stays the best vertex and is returned exactly."""
    c1, c2, chi2, scale = oracle.fit(theory3, Q3, Q3, theory3(1.0, 0.03))
    assert (c1, c2, chi2) == (1.0, 0.03, 0.0)
    assert scale == 1.0


def test_nelder_mead_stops_on_the_bound(oracle):
    """A target outside the box: every trial vertex is clamped, so the fit
    ends exactly on the bound it wanted to cross."""
    c1, c2, chi2, _ = oracle.fit(theory3, Q3, Q3, [1.0, 3.0, 0.4])
    assert c1 == 2.0
    assert 0.0 <= c2 <= 0.1 and chi2 > 0.0
    c1, c2, _, _ = oracle.fit(theory3, Q3, Q3, [1.0, 1.25, -0.3])
    assert c2 == 0.0 and abs(c1 - 1.25) < 1e-3


def test_grid_fit_equals_brute_force_scan(oracle):
    """The first strict minimum of the row-major (c1 outer, c2 inner) scan, as
    src/fit.rs:446-516 takes it, against a numpy scan with lstsq chi2."""
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(5)
    tq = np.linspace(0.0, 0.5, 50)
    prof = np.stack([1e4 * np.exp(-20 * tq ** 2), 3e3 * np.exp(-25 * tq ** 2),
                     50 + 0 * tq, 4e3 * np.exp(-22 * tq ** 2), 40 * np.exp(-5 * tq ** 2),
                     20 * np.exp(-8 * tq ** 2)])

    def total(c1, c2):
        AA, EE, HH, AE, AH, EH = prof
        return AA + c1 * c1 * EE + c2 * c2 * HH - 2 * c1 * AE + 2 * c2 * AH - 2 * c1 * c2 * EH

    q = np.sort(rng.uniform(0.02, 0.48, 30))
    inten = CubicSpline(tq, total(1.1, 0.05), bc_type="natural")(q) * (1 + 0.01 * rng.normal(size=30))
    sigma = 0.01 * np.abs(inten) + 1.0
    n1, n2 = 13, 9
    best = None
    for i in range(n1):
        c1 = 0.5 + (2.0 - 0.5) * i / (n1 - 1)
        for j in range(n2):
            c2 = 0.0 + (0.1 - 0.0) * j / (n2 - 1)
            ith = CubicSpline(tq, total(c1, c2), bc_type="natural")(q)
            _, res, _, _ = np.linalg.lstsq((ith / sigma)[:, None], inten / sigma, rcond=None)
            chi = res[0] / len(q)
            if best is None or chi < best[2] * (1 - 1e-9):
                best = (c1, c2, chi)
    c1, c2, chi2, _ = oracle.fit(total, tq, q, inten, sigma, method="grid", c1_points=n1,
                                 c2_points=n2)
    assert (c1, c2) == best[:2]
    assert abs(chi2 - best[2]) <= 1e-8 * best[2]

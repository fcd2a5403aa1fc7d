"""Validation helpers (src/validate.jl) against independent computations:
the textbook KS statistic (scipy) where no values tie, a brute-force
restatement of the reference's tie-order walk on tiny tied samples, hand
values, and exact power laws for the convergence rate."""
import itertools

import numpy as np
import pytest
from scipy import stats

import scrna_amd as sc


def walk_brute(x, y):
    """The reference's loop (validate.jl:24-39), written out literally."""
    sx, sy = sorted(x), sorted(y)
    nx, ny = len(sx), len(sy)
    ix = iy = 0
    d = ex = ey = 0.0
    while ix < nx or iy < ny:
        if ix >= nx:
            ey += 1.0 / ny
            iy += 1
        elif iy >= ny:
            ex += 1.0 / nx
            ix += 1
        elif sx[ix] <= sy[iy]:
            ex += 1.0 / nx
            ix += 1
        else:
            ey += 1.0 / ny
            iy += 1
        d = max(d, abs(ex - ey))
    return d


def test_ks_matches_scipy_without_ties():
    rng = np.random.default_rng(0)
    for nx, ny in [(50, 70), (500, 333), (1, 9)]:
        x, y = rng.normal(size=nx), rng.normal(0.3, 1.2, size=ny)
        assert sc.ks_statistic(x, y) == pytest.approx(stats.ks_2samp(x, y).statistic, abs=1e-12)


def test_ks_tie_order_is_the_references():
    # With ties the walk advances x first and measures after every step, so
    # identical samples of counts give 1/n instead of 0.
    assert sc.ks_statistic([1, 2, 3], [1, 2, 3]) == pytest.approx(1.0 / 3.0)
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = rng.integers(0, 4, size=rng.integers(1, 7)).tolist()
        y = rng.integers(0, 4, size=rng.integers(1, 7)).tolist()
        assert sc.ks_statistic(x, y) == walk_brute(x, y)            # bit for bit


def test_ks_all_small_permutations():
    for x in itertools.product(range(3), repeat=3):
        for y in itertools.product(range(3), repeat=2):
            assert sc.ks_statistic(x, y) == walk_brute(list(x), list(y))


def test_ks_critical_value():
    assert sc.ks_critical_value(100, 100) == pytest.approx(1.36 * np.sqrt(0.02))
    assert sc.ks_critical_value(100, 400, alpha=0.01) == pytest.approx(1.63 * np.sqrt(500 / 40000))
    assert sc.ks_critical_value(10, 10, alpha=0.1) == pytest.approx(1.22 * np.sqrt(0.2))


def test_compare_moments():
    r = sc.compare_moments([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])
    assert r == (0.0, 0.0, True)
    r = sc.compare_moments([1.0, 2.0, 3.0, 4.0], [2.0, 4.0, 6.0, 8.0])
    assert r.mean_relerr == pytest.approx(0.5) and r.var_relerr == pytest.approx(0.75) and not r.passed
    r = sc.compare_moments([0.0, 0.0], [0.0, 0.0])
    assert r.mean_relerr == 0.0 and r.var_relerr == 0.0 and r.passed


@pytest.mark.parametrize("rate", [1.0, 2.0, 0.5])
def test_convergence_rate_exact_power_law(rate):
    dts = np.array([0.1, 0.05, 0.02, 0.01])
    assert sc.convergence_rate(dts, 3.0 * dts ** rate) == pytest.approx(rate, abs=1e-12)

"""The estimators' error analysis (safire_amd.stats) against the properties the reference's
result tools are tested for (python/tests/safiretools/results/test_results_stats.py): bins
average consecutive samples, an AR(1) series recovers its correlation time and statistical
inefficiency (1 + rho) / (1 - rho), uncorrelated and constant series are not rebinned, the
slowest component sets the bin size, and the jackknife of a linear function is the plain
mean and standard error."""
import numpy as np
import pytest

from safire_amd import stats


def ar1(n, rho, rng):
    x = np.empty(n)
    x[0] = rng.normal()
    e = rng.normal(size=n) * np.sqrt(1 - rho * rho)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + e[i]
    return x


def test_bins_average_consecutive_samples():
    bins, _ = stats.rebinning_analysis(np.arange(13.0), rebinsize=4)
    np.testing.assert_allclose(bins, [1.5, 5.5, 9.5])
    bins, _ = stats.rebinning_analysis(np.arange(12.0), rebinsize=4, skip=4)
    np.testing.assert_allclose(bins, [5.5, 9.5])
    x = np.arange(24.0).reshape(12, 2)
    bins, _ = stats.rebinning_analysis(x, rebinsize=4)
    assert bins.shape == (3, 2)
    np.testing.assert_allclose(bins[:, 0], [3.0, 11.0, 19.0])
    with pytest.raises(ValueError):
        stats.rebinning_analysis(np.arange(7.0), rebinsize=4)


def test_uncorrelated_and_correlated_series():
    rng = np.random.default_rng(1)
    white = rng.normal(size=20000)
    bins, tau = stats.rebinning_analysis(white, rebinsize=stats.optimal_rebinsize(white))
    assert tau == pytest.approx(0.0, abs=0.5)
    assert stats.standard_error(bins) == pytest.approx(1 / np.sqrt(white.size), rel=0.25)
    assert stats.optimal_rebinsize(white) <= 4
    rho = 0.9
    x = ar1(100000, rho, np.random.default_rng(2))
    B = stats.optimal_rebinsize(x)
    assert B > 10 * rho / (1 - rho)
    bins, tau = stats.rebinning_analysis(x, rebinsize=B)
    assert tau == pytest.approx(rho / (1 - rho), rel=0.3)
    assert stats.standard_error(bins) > 3 * np.std(x) / np.sqrt(x.size)
    k = stats.statistical_inefficiency(ar1(200000, 0.8, np.random.default_rng(3)))
    assert k == pytest.approx((1 + 0.8) / (1 - 0.8), rel=0.2)
    assert stats.statistical_inefficiency(rng.normal(size=20000)) == pytest.approx(1.0, abs=0.1)


def test_degenerate_series():
    c = np.full((400, 4), 2.5)
    B = stats.optimal_rebinsize(c)
    bins, tau = stats.rebinning_analysis(c, rebinsize=B)
    assert B == 1 and tau == 0.0
    np.testing.assert_array_equal(stats.standard_error(bins), np.zeros(4))
    assert stats.statistical_inefficiency(np.full(50, 3.0)) == 1.0
    assert stats.statistical_inefficiency(1.0 + 1e-15 * np.arange(50.0)) == 1.0
    assert stats.statistical_inefficiency(np.ones(1)) == 1.0
    rng = np.random.default_rng(5)
    assert min(stats.statistical_inefficiency(rng.normal(size=60)) for _ in range(50)) == 1.0
    assert 100 // stats.optimal_rebinsize(np.arange(100.0)) > 2
    assert 180 // stats.optimal_rebinsize(ar1(180, 0.9, np.random.default_rng(6))) >= 5


def test_slowest_component_sets_the_bin_size():
    fast = ar1(20000, 0.3, np.random.default_rng(7))
    slow = ar1(20000, 0.95, np.random.default_rng(8))
    assert stats.optimal_rebinsize(slow) > stats.optimal_rebinsize(fast)
    assert stats.optimal_rebinsize(np.column_stack([fast, slow])) == stats.optimal_rebinsize(slow)
    assert stats.optimal_rebinsize(fast + 1j * slow) == stats.optimal_rebinsize(slow)
    x = np.concatenate([np.full(1000, 50.0), slow])
    assert stats.optimal_rebinsize(x, skip=1000) == stats.optimal_rebinsize(slow)


def test_jackknife_of_a_linear_function_is_the_mean():
    rng = np.random.default_rng(9)
    a, b = rng.normal(size=(40, 3)), rng.normal(size=(40, 3)) + 1j * rng.normal(size=(40, 3))
    v, e = stats.jackknife(lambda x, y: 2.0 * x - y, a, b)
    np.testing.assert_allclose(v, 2 * a.mean(0) - b.mean(0), rtol=1e-12)
    np.testing.assert_allclose(e, stats.standard_error(2 * a - b), rtol=1e-10)
    # a ratio: the jackknife removes the O(1/m) bias of the plug-in estimate
    x = 1.0 + 0.3 * rng.normal(size=200)
    y = 2.0 + 0.3 * rng.normal(size=200)
    v, e = stats.jackknife(lambda p, q: p / q, x, y)
    assert abs(v - 0.5) < 4 * e
    with pytest.raises(ValueError):
        stats.jackknife(lambda p, q: p, x, y[:10])

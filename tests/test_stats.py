"""Host statistics (functions/statistics.h:30-51) and binning."""
import numpy as np

from qcd_2d_gpu_amd import stats


def test_mean_variance_population():
    x = np.array([1.0, 2.0, 4.0, 7.0])
    m, v = stats.mean_variance(x)
    assert m == 3.5 and abs(v - np.var(x)) < 1e-15
    assert abs(stats.naive_error(x) - np.sqrt(np.var(x) / 3)) < 1e-15


def test_autocorrelation_of_ar1():
    rng = np.random.default_rng(1)
    phi, n = 0.7, 200000
    x = np.empty(n)
    x[0] = 0
    eps = rng.normal(size=n)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + eps[i]
    m, v = stats.mean_variance(x)
    rho = stats.autocorrelation(x, 5, m, v)
    assert np.allclose(rho, phi ** np.arange(1, 6), atol=0.02)
    tau = stats.integrated_autocorrelation_time(x)
    assert abs(tau - (0.5 + phi / (1 - phi))) < 0.3


def test_binned_error_of_independent_data():
    rng = np.random.default_rng(2)
    x = rng.normal(size=10000)
    m, e = stats.binned_error(x, 20)
    assert abs(e - 1 / np.sqrt(10000)) < 0.004
    assert abs(m - x[: 10000 // 20 * 20].mean()) < 1e-15


def test_autocorrelation_lags_beyond_series_are_zero():
    x = np.arange(5.0)
    m, v = stats.mean_variance(x)
    rho = stats.autocorrelation(x, 8, m, v)
    assert np.all(rho[5:] == 0.0)

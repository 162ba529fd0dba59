"""Host statistics of a measurement series (functions/statistics.h:30-51),
plus the binning the reference's README recommends for final errors
(README.md:74)."""
from __future__ import annotations

import numpy as np


def mean_variance(m) -> tuple[float, float]:
    """statistics.h:30-39: mean and population variance s2/N - mean^2; the
    printed error is sqrt(var / (N - 1))."""
    a = np.asarray(m, dtype=np.float64)
    n = a.size
    s = 0.0
    s2 = 0.0
    for v in a:  # sequential sums, the reference's order
        s += v
        s2 += v * v
    mean = s / n
    return mean, s2 / n - mean * mean


def naive_error(m) -> float:
    n = len(m)
    return float(np.sqrt(mean_variance(m)[1] / (n - 1)))


def autocorrelation(m, W: int, mean: float, var: float) -> np.ndarray:
    """statistics.h:41-51: C(i) = sum_j (m_j - mean)(m_{j+i} - mean) / (N - i)
    / var for i = 1..W (lags with no pairs give 0)."""
    a = np.asarray(m, dtype=np.float64) - mean
    n = a.size
    out = np.zeros(W)
    for i in range(1, W + 1):
        if n - i > 0 and var != 0.0:
            out[i - 1] = float(np.dot(a[: n - i], a[i:])) / (n - i) / var
    return out


def binned_error(m, nbins: int) -> tuple[float, float]:
    """Mean and the standard error from nbins equal blocks (trailing
    measurements that do not fill a block are dropped)."""
    a = np.asarray(m, dtype=np.float64)
    b = a.size // nbins
    if b < 1 or nbins < 2:
        raise ValueError("not enough measurements for the bins")
    blocks = a[: b * nbins].reshape(nbins, b).mean(axis=1)
    return float(blocks.mean()), float(blocks.std(ddof=1) / np.sqrt(nbins))


def integrated_autocorrelation_time(m, window: int | None = None) -> float:
    """tau_int = 1/2 + sum_{t=1}^{W} rho(t) with a self-consistent window
    (stop at the first W >= 5 tau_int)."""
    a = np.asarray(m, dtype=np.float64)
    mean, var = float(a.mean()), float(a.var())
    if var == 0.0:
        return 0.5
    wmax = window if window is not None else max(1, a.size // 4)
    rho = autocorrelation(a, wmax, mean, var)
    tau = 0.5
    for w in range(1, wmax + 1):
        tau += rho[w - 1]
        if window is None and w >= 5 * tau:
            break
    return float(tau)

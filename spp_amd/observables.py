"""Time-series statistics of the sweep driver: the reference's autocorrelation error bars.

* ``stat_error`` restates Stat<T>::Compute and Find_Correlation_Length (analyze/statistics.h:35-82)
  with the autocovariance of every lag from one FFT (numpy), instead of the O(n^2) double loop.
* ``polarity_error`` restates polarization-analysis.py:20-62 (drop the first half of the series,
  coarse lags, error = sqrt(var / (n // ic))); where the reference script would raise (no lag
  below 0.01) it returns an infinite error.
Both return (mean, error, correlation length in samples).
"""
from __future__ import annotations

import math

import numpy as np


def _autocov(x, mean):
    """sum_{i < n - tau} (x_i - m)(x_{i + tau} - m) for every tau (zero-padded FFT)."""
    d = np.asarray(x, dtype=np.float64) - mean
    n = d.size
    size = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(d, size)
    return np.fft.irfft(f * np.conj(f), size)[:n]


def stat_error(data):
    x = np.asarray(data, dtype=np.float64)
    n = x.size
    mean = float(x.sum() / n)
    var = float(np.dot(x, x) / n - mean * mean)
    half = n // 2
    # the reference's E[x^2] - E[x]^2 can round to a tiny negative or positive value for a
    # constant series; such a series has no fluctuation and no error
    if half == 0 or var <= 1e-300 or var <= 1e-14 * mean * mean:
        return mean, 0.0, max(half - 1, 0)
    c = _autocov(x, mean)[:half] / (n - np.arange(half)) / var
    below = np.nonzero(c < 0.01)[0]
    corr_len = int(below[0]) if below.size else half - 1
    return mean, math.sqrt(var * corr_len / n), corr_len


def polarity_error(t, p):
    p = np.asarray(p, dtype=np.float64)
    p = p[p.size // 2:]
    n = p.size
    m = n // 2
    di = max(m // 20, 1)
    ave = float(p.mean())
    var = float(p.var())
    if m == 0 or var <= 1e-300 or var <= 1e-14 * ave * ave:
        return ave, 0.0, None
    lags = np.arange(0, m, di)
    c = _autocov(p, ave)[lags] / (n - lags) / var
    idx = np.nonzero(c <= 0.01)[0]
    if idx.size == 0:
        return ave, math.inf, None
    ic = int(idx[0]) * di
    return ave, math.sqrt(var / (n // ic)), ic

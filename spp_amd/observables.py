"""Time-series statistics of the sweep driver: the reference's autocorrelation error bars.

* ``stat_error`` restates Stat<T>::Compute and Find_Correlation_Length (analyze/statistics.h:35-82)
  with the autocovariance of every lag from one FFT (numpy), instead of the O(n^2) double loop.
* ``polarity_error`` restates polarization-analysis.py:20-62 (drop the first half of the series,
  coarse lags, error = sqrt(var / (n // ic))); where the reference script would raise (no lag
  below 0.01) it returns an infinite error.
Both return (mean, error, correlation length in samples).
* ``order_parameters`` and ``quantities_time`` restate the trajectory analyses of
  analyze/analyze.h:71-148 over frames of headings and positions (host numpy).
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


def _polar_per_frame(theta):
    th = np.asarray(theta, dtype=np.float64)
    return np.hypot(np.cos(th).mean(axis=-1), np.sin(th).mean(axis=-1))


def order_parameters(theta_frames, Lx, Ly):
    """Compute_Order_Parameters (analyze/analyze.h:127-148): for frames of headings (shape
    [frames, particles]) and the box half-widths, returns (mean polarisation, its autocorrelation
    error, sigma2 = var(p) 4 Lx Ly, G = 1 - <p^4> / (3 <p^2>)).  G keeps the reference's
    normalisation: the Binder cumulant would divide by 3 <p^2>^2."""
    p = _polar_per_frame(theta_frames)
    mean, err, _ = stat_error(p)
    mean_square = float(np.dot(p, p) / p.size)
    var = mean_square - mean * mean
    p4 = float(np.mean(p ** 4))
    return mean, err, var * 4.0 * Lx * Ly, 1.0 - p4 / (3.0 * mean_square)


def quantities_time(theta_frames, x_frames, y_frames, start=0):
    """Quantities_Time (analyze/analyze.h:71-100): for every frame from `start` on, the
    polarisation |<(cos, sin)>|, the nematic order S = |<(cos 2 theta, sin 2 theta)>| and the mean
    squared displacement from frame `start` (positions as the frames store them: unwrapped for the
    repulsive and active Brownian models).  Returns three arrays."""
    th = np.asarray(theta_frames, dtype=np.float64)[start:]
    c, s_ = np.cos(th), np.sin(th)
    p = np.hypot(c.mean(axis=1), s_.mean(axis=1))
    c2 = (2 * c * c - 1).mean(axis=1)
    s2 = (2 * c * s_).mean(axis=1)
    S = np.sqrt(c2 * c2 + s2 * s2)
    x = np.asarray(x_frames, dtype=np.float64)
    y = np.asarray(y_frames, dtype=np.float64)
    dr2 = ((x[start:] - x[start]) ** 2 + (y[start:] - y[start]) ** 2).mean(axis=1)
    return p, S, dr2

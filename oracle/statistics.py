"""Loop restatements of the reference's time-series error estimators (TEST INFRASTRUCTURE ONLY).

* stat_error: Stat<T>::Compute + Find_Correlation_Length (analyze/statistics.h:35-82): population
  variance, correlation length = first lag tau < n/2 whose normalised autocovariance
  sum_{i<n-tau} (x_i - m)(x_{i+tau} - m)/(n - tau) / var drops below 0.01 (else n/2 - 1),
  error = sqrt(var * tau / n).
* polarity_error: polarization-analysis.py:20-62 (Python 2 integer division restated): drop the
  first half, autocorrelation at lags 0, di, 2 di, ... below m = n/2 with di = m/20, ic = di times
  the index of the first value <= 0.01, error = sqrt(var / (n // ic)).
"""
from __future__ import annotations

import math


def stat_error(data):
    n = len(data)
    mean = sum(data) / n
    var = sum(x * x for x in data) / n - mean * mean
    corr_len = 0
    tau = 0
    while tau < n // 2:
        c = 0.0
        for i in range(n - tau):
            c += (data[i] - mean) * (data[i + tau] - mean) / (n - tau)
        c = c / var if var != 0 else float("nan")
        corr_len = tau
        if c < 0.01:
            break
        tau += 1
    return mean, math.sqrt(var * corr_len / n), corr_len


def polarity_error(t, p):
    n0 = len(p)
    t = list(t)[n0 // 2:]
    p = list(p)[n0 // 2:]
    n = len(p)
    m = n // 2
    di = max(m // 20, 1)
    ave = sum(p) / n
    var = sum((x - ave) ** 2 for x in p) / n
    cl = []
    for i in range(0, m, di):
        ci = 0.0
        for j in range(n - i):
            ci += (p[j] - ave) * (p[j + i] - ave) / (n - i)
        cl.append(ci / var)
    ic = 0
    while ic < len(cl) and cl[ic] > 0.01:
        ic += 1
    if ic == len(cl):
        return ave, float("nan"), None
    ic *= di
    neff = n // ic
    return ave, math.sqrt(var / neff), ic

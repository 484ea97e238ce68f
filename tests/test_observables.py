"""The sweep driver's error-bar estimators against the loop restatements of the reference
(oracle/statistics.py) and against what an AR(1) series fixes in closed form."""
import math

import numpy as np
import pytest

from spp_amd import observables as OB


def ar1(phi, n, seed, sigma=1.0):
    from scipy.signal import lfilter

    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n) * sigma
    e[0] /= math.sqrt(1 - phi * phi)  # stationary start
    return lfilter([1.0], [1.0, -phi], e)


@pytest.mark.parametrize("phi,n", [(0.0, 400), (0.5, 500), (0.9, 700), (0.98, 600)])
def test_stat_error_equals_reference_loops(oracle_mod, phi, n):
    import oracle.statistics as OS

    x = 0.5 + 0.1 * ar1(phi, n, 3)
    m, e, c = OB.stat_error(x)
    wm, we, wc = OS.stat_error(list(x))
    assert c == wc
    assert abs(m - wm) < 1e-12 and abs(e - we) < 1e-12 * max(1.0, we)


@pytest.mark.parametrize("phi,n", [(0.3, 800), (0.8, 900), (0.95, 1200)])
def test_polarity_error_equals_reference_script(oracle_mod, phi, n):
    import oracle.statistics as OS

    x = 0.3 + 0.05 * ar1(phi, n, 5)
    t = np.arange(n) * 0.5
    m, e, ic = OB.polarity_error(t, x)
    wm, we, wic = OS.polarity_error(list(t), list(x))
    assert ic == wic
    assert abs(m - wm) < 1e-12 and abs(e - we) < 1e-12


def test_correlation_length_of_ar1():
    """AR(1) autocorrelation phi^tau falls below 0.01 at tau = ceil(log 0.01 / log phi)."""
    for phi in (0.5, 0.8, 0.9):
        x = ar1(phi, 2_000_000, 7)
        _, _, c = OB.stat_error(x)
        # the sample autocorrelation has noise ~1e-3 at this length: the first crossing lies
        # where phi^tau is between 0.02 and 0.004
        lo = math.log(0.02) / math.log(phi) - 1
        hi = math.log(0.004) / math.log(phi) + 1
        assert lo <= c <= hi, (phi, c, lo, hi)


def test_degenerate_series():
    m, e, c = OB.stat_error(np.full(50, 0.7))
    assert m == pytest.approx(0.7) and e == 0.0
    m, e, ic = OB.polarity_error(np.arange(50), np.full(50, 0.7))
    assert e == 0.0

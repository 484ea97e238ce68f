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


# ---------------------------------------------------------------------------------------------
# Trajectory analyses (analyze/analyze.h:71-148), pinned by closed forms
# ---------------------------------------------------------------------------------------------

def test_order_parameters_closed_forms():
    from spp_amd.observables import order_parameters

    n = 500
    aligned = np.full((40, n), 0.7)
    m, err, sigma2, G = order_parameters(aligned, 10.0, 5.0)
    assert m == pytest.approx(1.0, abs=1e-15) and err == 0.0 and abs(sigma2) < 1e-12
    assert G == pytest.approx(2.0 / 3.0, abs=1e-15)  # 1 - 1/(3*1): the reference's normalisation
    # frames alternating between polarisation a (a fraction of the headings at 0, the rest split
    # at +-pi/2) and 1: closed-form moments
    k = 300
    f = np.zeros(n)
    f[k:k + (n - k) // 2] = np.pi / 2
    f[k + (n - k) // 2:] = -np.pi / 2
    a = k / n
    frames = np.array([f if i % 2 else np.zeros(n) for i in range(40)])
    m, err, sigma2, G = order_parameters(frames, 10.0, 5.0)
    assert m == pytest.approx((a + 1) / 2, rel=1e-13)
    var = (a * a + 1) / 2 - ((a + 1) / 2) ** 2
    assert sigma2 == pytest.approx(var * 4 * 10.0 * 5.0, rel=1e-10)
    assert G == pytest.approx(1 - ((a ** 4 + 1) / 2) / (3 * (a * a + 1) / 2), rel=1e-13)


def test_quantities_time_closed_forms():
    from spp_amd.observables import quantities_time

    rng = np.random.default_rng(3)
    n, T = 400, 5
    x0, y0 = rng.uniform(-5, 5, n), rng.uniform(-5, 5, n)
    d = np.array([[0.0, 0.0], [0.3, -0.4], [1.0, 2.0], [-3.0, 0.0], [0.0, 0.5]])
    x = np.array([x0 + v[0] for v in d])
    y = np.array([y0 + v[1] for v in d])
    th = np.empty((T, n))
    th[0] = 0.3                                  # aligned: p = 1, S = 1
    th[1] = np.where(np.arange(n) % 2, 0.3, 0.3 + np.pi)  # antiparallel: p = 0, S = 1
    th[2] = np.where(np.arange(n) % 2, 0.0, np.pi / 2)    # perpendicular pairs: S = 0, p = 1/sqrt 2
    th[3] = 2 * np.pi * np.arange(n) / n          # uniform: p = S = 0
    th[4] = 1.0
    p, S, dr2 = quantities_time(th, x, y)
    assert p == pytest.approx([1, 0, 2 ** -0.5, 0, 1], abs=1e-12)
    assert S == pytest.approx([1, 1, 0, 0, 1], abs=1e-12)
    assert dr2 == pytest.approx((d ** 2).sum(axis=1), rel=1e-12, abs=1e-15)
    p1, _, dr1 = quantities_time(th, x, y, start=1)
    assert dr1 == pytest.approx(((d[1:] - d[1]) ** 2).sum(axis=1), rel=1e-12, abs=1e-15)

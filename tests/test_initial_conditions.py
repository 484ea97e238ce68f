"""Initial-condition builders of the C ABI (host functions; no GPU): closed forms and geometry."""
import math

import numpy as np
import pytest

import spp_amd as S


def test_triangle_lattice_known_case():
    # n = 4 in a 4 x 4 box: Nx = (int)sqrt(2*4*sqrt(3)/4) + 1 = 2, spacing 2 (set-up.h:116-145)
    x, y = S.triangle_lattice(4, 2.0, 2.0, 0.0)
    h = math.sqrt(3.0)
    want = [(-2.0, -2.0), (0.0, -2.0), (-1.0, h - 2.0), (1.0, h - 2.0)]
    assert np.allclose(np.stack([x, y], 1), want, atol=1e-15)


def test_triangle_lattice_is_triangular():
    n, L, sigma = 900, 16.0, 0.5
    x, y = S.triangle_lattice(n, L, L, sigma)
    assert np.all(np.abs(x) <= L) and np.all(np.abs(y) <= L)
    nx = int(math.sqrt(L * n * math.sqrt(3) / (2 * L))) + 1
    s = (2 * L - sigma) / nx
    d = np.hypot(x[:, None] - x[None], y[:, None] - y[None])
    np.fill_diagonal(d, np.inf)
    # interior sites have six neighbours at the lattice spacing
    inner = (np.abs(x) < L - 2 * s) & (np.abs(y) < L - 2 * s)
    near = (np.abs(d - s) < 1e-9 * s).sum(axis=1)
    assert np.all(d.min(axis=1) > s * (1 - 1e-9))
    assert np.all(near[inner] == 6)


def test_ring_membrane():
    nm = 1024
    x, y = S.ring_membrane(nm, bead_diameter=1.0)
    R = 0.5 / math.sin(math.pi / nm)
    assert np.allclose(np.hypot(x, y), R, rtol=1e-13)
    step = np.hypot(np.diff(np.append(x, x[0])), np.diff(np.append(y, y[0])))
    assert np.allclose(step, 1.0, rtol=1e-9)
    x2, y2 = S.ring_membrane(nm, radius=180.0)  # a given radius: bonds stretched to 1.1045
    assert np.allclose(np.hypot(x2, y2), 180.0, rtol=1e-13)
    assert np.allclose(np.hypot(x2[1] - x2[0], y2[1] - y2[0]), 360 * math.sin(math.pi / nm), rtol=1e-12)


def test_disk_state_is_uniform():
    from scipy import stats

    n, R = 20000, 180.0
    x, y, th = S.disk_state(3, n, R)
    r = np.hypot(x, y)
    assert r.max() <= R
    assert stats.kstest((r / R) ** 2, "uniform").pvalue > 1e-3
    assert stats.kstest(np.arctan2(y, x) / (2 * np.pi) + 0.5, "uniform").pvalue > 1e-3
    assert stats.kstest(th / (2 * np.pi), "uniform").pvalue > 1e-3
    x2, _, _ = S.disk_state(3, n, R)
    assert np.array_equal(x, x2)


def test_builders_reject_bad_input():
    with pytest.raises(S.SppError):
        S.ring_membrane(2)
    with pytest.raises(S.SppError):
        S.disk_state(1, 10, 0.0)

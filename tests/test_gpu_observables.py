"""GPU: the isotropic pair distribution g(r) from the Verlet list against brute force."""
import numpy as np
import pytest

import spp_amd as S
from tests.helpers import c2_params

pytestmark = pytest.mark.gpu


def brute_counts(x, y, L, r_max, nbins):
    """Ordered-pair distance histogram by brute force (minimum image)."""
    dr = r_max / nbins
    out = np.zeros(nbins, dtype=np.int64)
    for i in range(x.size):
        dx = x[i] - x
        dy = y[i] - y
        dx -= 2 * L * np.rint(dx / (2 * L))
        dy -= 2 * L * np.rint(dy / (2 * L))
        d = np.sqrt(dx * dx + dy * dy)
        d[i] = np.inf
        b = (d[d < r_max] / dr).astype(np.int64)
        np.add.at(out, b[b < nbins], 1)
    return out


@pytest.mark.parametrize("n,L,r_max,nbins", [(3000, 24.0, 1.3, 26), (5000, 20.0, 1.0, 50)])
def test_pair_counts_equal_brute_force(oracle_mod, n, L, r_max, nbins):
    x, y, th = oracle_mod.random_state(3, n, L, L)
    e = S.Engine(c2_params(n=n, L=L))
    e.set_state(x, y, th)
    r, g, counts = e.pair_distribution(r_max, nbins)
    want = brute_counts(x, y, L, r_max, nbins)
    assert np.abs(counts - want).sum() <= 2, np.nonzero(counts != want)  # bin-edge roundings at most
    # an ideal-gas state: g = 1 within counting noise
    err = np.sqrt(np.maximum(want, 1)) / np.maximum(want, 1)
    assert np.all(np.abs(g - 1) < 5 * err + 1e-12)


def test_pair_distribution_rejects_beyond_list():
    x, y, th = S.random_state(1, 1000, 16.0, 16.0)
    e = S.Engine(c2_params(n=1000, L=16.0))
    e.set_state(x, y, th)
    with pytest.raises(S.SppError):
        e.pair_distribution(2.0, 10)

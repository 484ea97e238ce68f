"""The allocation-free pair distribution (spp_pair_distribution_noalloc) against the allocating
one (spp_pair_distribution): bit-for-bit equal counts and g on two calls in a row, the second on a
smaller engine with fewer bins (the shared static scratch keeps the first call's contents, which
must not leak into the second), the size limit, and the state checks."""
import numpy as np
import pytest

import spp_amd as S
from tests.helpers import c2_params

pytestmark = pytest.mark.gpu


def _engine(n, L, seed):
    e = S.Engine(c2_params(n=n, L=L))
    x, y, th = S.random_state(seed, n, L, L)
    e.set_state(x, y, th)
    e.multi_step(20)
    return e


@pytest.mark.parametrize("with_counts", [True, False])
def test_noalloc_pair_distribution_equals_allocating_one(with_counts):
    cases = [(_engine(16384, 64.0, 3), 1.3, 4000), (_engine(4096, 32.0, 4), 1.0, 700)]
    for e, r_max, nbins in cases:  # the second call: smaller engine, fewer bins, same scratch
        _, g_ref, c_ref = e.pair_distribution(r_max, nbins)
        counts = np.full(nbins, -7, dtype=np.int64) if with_counts else None
        g = np.full(nbins, np.nan)
        e.pair_distribution_noalloc(r_max, nbins, counts=counts, g=g)
        if with_counts:
            assert np.array_equal(counts, c_ref)
        assert np.array_equal(g, g_ref)
        assert c_ref.sum() > 0


def test_noalloc_pair_distribution_limits():
    e = _engine(4096, 32.0, 5)
    with pytest.raises(S.SppError):
        e.pair_distribution_noalloc(1.0, 8193, g=np.empty(8193))
    counts = np.empty(8192, dtype=np.int64)
    e.pair_distribution_noalloc(1.3, 8192, counts=counts)  # the maximum itself
    assert np.array_equal(counts, e.pair_distribution(1.3, 8192)[2])
    with pytest.raises(S.SppError):
        e.pair_distribution_noalloc(2.0, 100, g=np.empty(100))  # beyond the list radius

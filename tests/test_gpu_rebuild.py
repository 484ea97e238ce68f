"""The two rebuild sorts: the one-launch tile sort (SPP_REBUILD_SORT_AUTO) against the general
counting sort (SPP_REBUILD_SORT_GENERAL).  Both order particles by (cell, caller index), so the
states they lead to are equal bit for bit; a tile sort that cannot work (a particle moved farther
than one cell) is detected and redone by the general sort, again with equal results."""
import numpy as np
import pytest

import spp_amd as S
from tests.helpers import c1_params, c2_params

pytestmark = pytest.mark.gpu


def _run(p, x, y, th, steps, sort):
    e = S.Engine(p.copy().update(rebuild_sort=sort))
    e.set_state(x, y, th)
    e.multi_step(steps)
    return e.get_state(), e.info(), e


@pytest.mark.parametrize("n,L,steps,D", [(16384, 64.0, 300, 0.05), (65536, 128.0, 200, 0.2),
                                          (262144, 256.0, 150, 0.05)])
def test_tile_sort_bitwise_equals_general(oracle_mod, n, L, steps, D):
    x, y, th = oracle_mod.random_state(41, n, L, L)
    p = c2_params(n=n, L=L, D=D)
    a, ia, ea = _run(p, x, y, th, steps, S.REBUILD_SORT_AUTO)
    b, ib, eb = _run(p, x, y, th, steps, S.REBUILD_SORT_GENERAL)
    assert ia.tile_rebuilds >= 3 and ia.tile_fallbacks == 0
    assert ib.tile_rebuilds == 0
    assert ia.rebuilds == ib.rebuilds
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    oa, ia_ = ea.get_neighbors()
    ob, ib_ = eb.get_neighbors()
    assert np.array_equal(oa, ob) and np.array_equal(ia_, ib_)


def test_tile_sort_vicsek_dense_cells(oracle_mod):
    """Continuous Vicsek at density 4 (about 7 particles per cell): tiles of few cells."""
    n, L = 16384, 32.0
    x, y, th = oracle_mod.random_state(43, n, L, L)
    p = c1_params(n=n, Lx=L, Ly=L, D_rot=0.05)
    a, ia, _ = _run(p, x, y, th, 200, S.REBUILD_SORT_AUTO)
    b, _, _ = _run(p, x, y, th, 200, S.REBUILD_SORT_GENERAL)
    assert ia.tile_rebuilds >= 2 and ia.tile_fallbacks == 0
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_tile_sort_falls_back_when_particles_jump_cells(oracle_mod):
    """Fixed-period rebuilds with a long period let particles travel several cells between builds:
    the tile sort misses them, flags the rebuild, and the general sort redoes it at the same point,
    so the trajectory still equals the general sort's bit for bit (three failures retire the tile
    path for the engine)."""
    n, L = 8192, 48.0
    x, y, th = oracle_mod.random_state(47, n, L, L)
    p = c2_params(n=n, L=L, dt=0.02, rebuild=S.REBUILD_PERIOD, rebuild_period=150, repulsion=S.REP_NONE)
    a, ia, _ = _run(p, x, y, th, 900, S.REBUILD_SORT_AUTO)
    b, _, _ = _run(p, x, y, th, 900, S.REBUILD_SORT_GENERAL)
    assert ia.tile_fallbacks >= 1
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_set_state_and_restore_after_tile_rebuilds(oracle_mod):
    """set_state hands over particles in caller order and restore a set sorted by an older build,
    so neither may take the tile sort (whose premise is the last build's order): after tile
    rebuilds, a restore replays bit for bit, and a fresh set_state on the used engine follows the
    oracle like a new engine does."""
    n, L = 16384, 64.0
    p = c2_params(n=n, L=L)
    x, y, th = oracle_mod.random_state(51, n, L, L)
    e = S.Engine(p)
    e.set_state(x, y, th)
    e.multi_step(100)
    assert e.info().tile_rebuilds > 0
    e.snapshot(1)
    e.multi_step(60)
    s1 = e.get_state()
    e.restore(1)
    e.multi_step(60)
    s2 = e.get_state()
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k
    x2, y2, th2 = oracle_mod.random_state(52, n, L, L)
    for _ in range(2):  # twice: the second upload replaces a state the tile sort produced
        e.set_state(x2, y2, th2)
        got = e.get_state()
        assert np.array_equal(got["xw"], x2) and np.array_equal(got["theta"], th2)
    e.multi_step(50)
    st = e.get_state()
    assert all(np.all(np.isfinite(v)) for v in st.values())
    q = p.copy().update(D_rot=0.0)  # noiseless: the used engine's step counter does not matter
    f = S.Engine(q)
    f.set_state(x2, y2, th2)
    f.multi_step(100)
    o = oracle_mod.Oracle(q)
    o.set_state(x2, y2, th2)
    o.multi_step(100)
    a, b = f.get_state(), o.get_state()
    assert np.abs(a["xu"] - b["xu"]).max() < 1e-9
    assert f.info().tile_rebuilds > 0 and f.info().tile_fallbacks == 0

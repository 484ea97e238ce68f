"""Deep overlaps: a step right after a rebuild moves a particle beyond the half skin (stiff
repulsion pushing apart a uniform random start).  The engine redoes that step on the wide list (a
5x5-cell stencil holding every pair up to two cell widths apart, whose skin is several times wider)
and returns to the ordinary list once the particles slow down, instead of ending the run.

The oracle runs the same physics with a list radius of 3.5 rebuilt after every step, so its pair
sets are exact for any displacement below 1.25 per step; the engine must follow it within the
trajectory tolerance.  Reference: the Verlet list of shared/cell.h:74-103 with rv = r_list."""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff, c2_params

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def _stiff(n=4096, L=32.0, **kw):
    # k = 100 over radius 1: overlapping pairs separate at up to ~100 per unit time, 0.5 per step
    return c2_params(n=n, L=L, A_rep=100.0, **kw)


def _oracle_exact(oracle_mod, p):
    q = p.copy().update(r_list=3.5, rebuild=S.REBUILD_PERIOD, rebuild_period=1)
    return oracle_mod.Oracle(q)


def _err(a, b):
    return max(np.abs(a["xu"] - b["xu"]).max(), np.abs(a["yu"] - b["yu"]).max(),
               angle_diff(a["theta"], b["theta"]).max())


def test_deep_overlap_runs_on_the_wide_list(oracle_mod):
    p = _stiff()
    x, y, th = oracle_mod.random_state(5, p.n, p.Lx, p.Ly)
    e = S.Engine(p)
    e.set_state(x, y, th)
    o = _oracle_exact(oracle_mod, p)
    o.set_state(x, y, th)
    e.multi_step(10)
    o.multi_step(10)
    info = e.info()
    assert info.wide_rebuilds > 0 and info.rollbacks > 0, (info.wide_rebuilds, info.rollbacks)
    err = _err(e.get_state(), o.get_state())
    assert err < TRAJ_TOL, err


def test_wide_list_is_left_once_particles_slow_down(oracle_mod):
    """After the overlaps have been pushed apart the engine rebuilds ordinary lists again, and the
    trajectory still follows the exact-pair oracle."""
    p = _stiff()
    x, y, th = oracle_mod.random_state(6, p.n, p.Lx, p.Ly)
    e = S.Engine(p)
    e.set_state(x, y, th)
    o = _oracle_exact(oracle_mod, p)
    o.set_state(x, y, th)
    e.multi_step(60)
    o.multi_step(60)
    w1, r1 = e.info().wide_rebuilds, e.info().rebuilds
    e.multi_step(60)
    o.multi_step(60)
    info = e.info()
    assert w1 > 0
    assert info.rebuilds > r1 and info.wide_rebuilds == w1  # the last 60 steps rebuilt ordinary lists
    assert _err(e.get_state(), o.get_state()) < TRAJ_TOL


def test_wide_list_holds_the_pairs_within_two_cell_widths(oracle_mod):
    """The wide list the first step ran on holds exactly the pairs closer than two cell widths at
    the positions it was built from (the uploaded start), checked by brute force."""
    p = _stiff(n=2048, L=32.0)
    x, y, th = oracle_mod.random_state(7, p.n, p.Lx, p.Ly)
    e = S.Engine(p)
    e.set_state(x, y, th)
    e.multi_step(1)  # the deep overlaps of this start put the first step on a wide list
    info = e.info()
    assert info.wide_rebuilds > 0
    off, idx = e.get_neighbors()
    n = p.n
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx -= 2 * p.Lx * np.rint(dx / (2 * p.Lx))
    dy -= 2 * p.Ly * np.rint(dy / (2 * p.Ly))
    d = np.sqrt(dx * dx + dy * dy)
    np.fill_diagonal(d, np.inf)
    r_wide = 2 * min(2 * p.Lx / info.ncell_x, 2 * p.Ly / info.ncell_y)
    got = np.zeros((n, n), dtype=bool)
    got[np.repeat(np.arange(n), np.diff(off)), idx] = True
    clear = np.abs(d - r_wide) > 1e-9
    assert np.array_equal(got[clear], (d < r_wide)[clear])

"""GPU: fixed-period intervals captured once as HIP graphs and replayed (graph_intervals = 1).

Every replay must give the states the launched kernels give, bitwise: a ragged batch and a single
box, RK2 and RK4, checked after every call with the program's own eager work (state reads,
polarisation, snapshots) and synchronisations between the replays; through a list-capacity
overflow (the buffers change, the graphs are dropped and captured again); and calls whose length
is not a multiple of the period (partial intervals run as plain launches).
"""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.test_gpu_batch import initial_states, ragged_systems, vicsek_params

pytestmark = pytest.mark.gpu


def pair(p, systems, x, y, th):
    a = S.Engine(p.copy().update(graph_intervals=1), systems)
    b = S.Engine(p.copy().update(graph_intervals=0), systems)
    a.set_state(x, y, th)
    b.set_state(x, y, th)
    return a, b


def same(a, b):
    sa, sb = a.get_state(), b.get_state()
    return all(np.array_equal(sa[k], sb[k]) for k in ("xw", "yw", "xu", "yu", "theta"))


@pytest.mark.parametrize("integrator", [S.RK2, S.RK4])
def test_batch_replays_bitwise(integrator):
    p = vicsek_params(integrator=integrator, rebuild=S.REBUILD_PERIOD, rebuild_period=20)
    systems = ragged_systems(5, seed=70 + integrator, n_hi=3000)
    x, y, th = initial_states(systems, 4)
    a, b = pair(p, systems, x, y, th)
    for call in (100, 60, 37, 100, 3, 100):
        a.multi_step(call)
        b.multi_step(call)
        assert same(a, b), call
        assert np.array_equal(a.batch_polarization(), b.batch_polarization())
    captures, replays, cached = a.graph_stats()
    # a list-capacity growth replaces the list buffers and drops the graphs held
    assert replays > captures > 0 and 0 < cached <= captures
    assert b.graph_stats() == (0, 0, 0)
    print("captures %d replays %d held %d, capacity %d" % (captures, replays, cached, a.info().ell_capacity))
    assert a.info().steps == b.info().steps == 400 and a.info().t == b.info().t


def test_single_box_snapshot_and_restore_between_replays():
    p = S.default_params(S.MODEL_REPULSIVE).update(
        n=4000, Lx=32.0, Ly=32.0, dt=0.005, v0=1.0, R_align=1.0, repulsion=S.REP_HARMONIC, A_rep=10.0,
        sigma_rep=1.0, rc_rep=1.0, D_rot=0.2, integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_PERIOD,
        rebuild_period=10, seed=8)
    x, y, th = S.random_state(2, p.n, p.Lx, p.Ly)
    a, b = pair(p, None, x, y, th)
    for e in (a, b):
        e.multi_step(50)
        e.snapshot(3)
        e.multi_step(70)
        e.restore(3)
        e.multi_step(80)
    assert same(a, b)
    assert a.graph_stats()[1] > 0


def test_capacity_overflow_drops_and_recaptures():
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=10, ell_capacity=12)
    systems = [S.system(n, 0.5 * math.sqrt(n / 4.0), v0=0.5, g=1.0, D_rot=0.2, seed=5 + n) for n in (600, 1500)]
    x, y, th = initial_states(systems, 6)
    a, b = pair(p, systems, x, y, th)
    for _ in range(4):
        a.multi_step(50)
        b.multi_step(50)
        assert same(a, b)
    assert a.info().ell_capacity > 12
    assert a.graph_stats()[1] > 0

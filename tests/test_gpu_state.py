"""The particle-level and state-vector surface of the boundary on the GPU:

* spp_reset_forces + spp_pair_interact over every listed pair (the reference's Particle::Reset and
  Particle::Interact, shared/particle.h:559-564, 728-757, 296-338, 964-1012) give spp_interact's
  forces and torques (the same pair terms summed in another order: within 1e-12 of the largest),
  for the repulsive, continuous (with the self term) and chain models and a batched engine;
* invalid pairs are rejected with an error and leave the forces alone;
* the C++ facade's Save(State) / Load(State) (parallel/box.h:222-267) replays a run exactly, a
  perturbed copy steps without error, and particle(i).Interact(particle(j)) over the list
  reproduces Box::Interact (state_demo, built with the engine).
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import spp_amd as S
from helpers import c1_params, c2_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "spp_amd", "state_demo")
TOL = 1e-12


def pairwise_forces(e):
    """reset + one pair_interact per unordered listed pair, in particle order."""
    off, idx = e.get_neighbors()
    e.reset_forces()
    n = 0
    for i in range(e.n):
        for j in idx[off[i]:off[i + 1]]:
            if j > i:
                e.pair_interact(i, int(j))
                n += 1
    assert n > 0
    return e.get_forces()


def check_same_forces(e):
    e.build_neighbors()
    e.interact()
    ref = e.get_forces()
    got = pairwise_forces(e)
    scale = max(np.abs(r).max() for r in ref)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= TOL * scale, np.abs(g - r).max() / scale
    return ref


@pytest.mark.parametrize("rep", [S.REP_HARMONIC, S.REP_R12_TRUNCATED, S.REP_YUKAWA_TRUNCATED])
def test_pair_interact_repulsive(rep):
    p = c2_params(n=1200, L=18.0, repulsion=rep, A_rep=2.0, rc_rep=1.05, R_align=1.1, r_list=1.4)
    e = S.Engine(p)
    e.set_state(*S.random_state(3, p.n, p.Lx, p.Ly))
    ref = check_same_forces(e)
    assert np.abs(ref[2]).max() > 0


def test_pair_interact_continuous_self_term():
    p = c1_params(n=1500, Lx=15.0, Ly=15.0, alpha=0.3)
    e = S.Engine(p)
    e.set_state(*S.random_state(4, p.n, p.Lx, p.Ly))
    check_same_forces(e)


@pytest.mark.parametrize("nb", [1, 3])
def test_pair_interact_chains(nb):
    """Swimmers only (ring bonds are not pair interactions), chains of nb beads."""
    p = S.default_params(S.MODEL_ABP)
    p.update(n=600, n_membrane=0, Lx=12.0, Ly=12.0, integrator=S.RK2, dt=1e-3, repulsion=S.REP_HARMONIC,
             A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, D_rot=0.1, v0=2.0, mob_par=0.5, mob_perp=0.5,
             mob_rot=1.0, seed=11, nb_swimmer=nb, bead_sep=1.0, sigma_bead=1.0, rc_bead=1.1, A_bead=1.0,
             r_list=1.1 + (nb - 1) * 1.0 + 0.4)
    e = S.Engine(p)
    e.set_state(*S.random_state(5, p.n, p.Lx, p.Ly))
    ref = check_same_forces(e)
    if nb > 1:
        assert np.abs(ref[2]).max() > 0  # bead-pair torques s x F


def test_pair_interact_batch_and_errors():
    p = c1_params()
    systems = [S.system(700, 9.0, v0=0.5, g=1.0, D_rot=0.1, seed=7),
               S.system(900, 11.0, v0=0.5, g=2.0, D_rot=0.2, seed=8)]
    e = S.Engine(p, systems)
    xs = [S.random_state(20 + k, s.n, s.Lx, s.Ly) for k, s in enumerate(systems)]
    e.set_state(*[np.concatenate([x[c] for x in xs]) for c in range(3)])
    check_same_forces(e)
    before = e.get_forces()
    for i, j in [(3, 3), (-1, 2), (0, e.n), (5, 800)]:  # same, out of range, two systems
        with pytest.raises(S.SppError):
            e.pair_interact(i, j)
    after = e.get_forces()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_state_demo_save_load_and_particles():
    assert os.path.exists(DEMO), "build() compiles spp_amd/state_demo"
    run = subprocess.run([DEMO, "3000", "28", "60"], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    out = run.stdout
    m = re.search(r"replay t=(\S+) (\S+) step=(\d+) (\d+) max_diff=(\S+) (\S+) (\S+)", out)
    assert m, out
    assert m.group(1) == m.group(2) and m.group(3) == m.group(4) == "120"
    assert all(float(v) == 0.0 for v in m.group(5, 6, 7)), out
    m = re.search(r"perturb initial=(\S+) final=(\S+)", out)
    assert m and abs(float(m.group(1)) - 1e-7) < 1e-12 and 0 < float(m.group(2)) < 1.0, out
    m = re.search(r"pairs (\d+) max_diff=(\S+) (\S+) (\S+) scale=(\S+)", out)
    assert m and int(m.group(1)) > 3000, out
    assert max(float(v) for v in m.group(2, 3, 4)) <= TOL * float(m.group(5)), out


def test_clock_round_trip_replays_bitwise():
    """get_state + get_clock at a rebuilt point, then set_state + set_clock: the next 150 steps
    equal the first run's bit for bit (same noise steps, same list rebuild steps)."""
    p = c2_params(n=4000, L=32.0)
    e = S.Engine(p)
    e.set_state(*S.random_state(9, p.n, p.Lx, p.Ly))
    e.multi_step(90)
    e.build_neighbors()
    st, clk = e.get_state(), e.get_clock()
    assert clk.step == 90 and max(clk.rebuild_estimate) > 0
    e.multi_step(150)
    a = e.get_state()
    e.multi_step(40)  # the engine's schedule moves on
    e.set_state(st["xw"], st["yw"], st["theta"], st["xu"], st["yu"])
    e.set_clock(clk)
    e.multi_step(150)
    b = e.get_state()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert e.get_clock().step == 240
    bad = S.Clock()
    bad.step = -1
    with pytest.raises(S.SppError):
        e.set_clock(bad)
    bad.step, bad.rebuild_estimate[1] = 3, float("nan")
    with pytest.raises(S.SppError):
        e.set_clock(bad)

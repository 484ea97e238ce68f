"""GPU: the persistent kernel (kernel_variant 3): one workgroup per system runs a whole
spp_multi_step in one launch, rebuilding from inside the kernel.

* Fixed rebuild period: the rebuild steps are those of the multi-launch engine, so every state is
  bitwise equal to it -- batched (ragged systems up to 4096 particles, RK2, RK4, Euler), single
  box, through a list-capacity overflow, and the polarisation samples recorded in the kernel.
* Displacement-triggered rebuilds (each system on its own): 100-step trajectories against an
  independent oracle run per system within 1e-9, and the same as the multi-launch engine to
  rounding.
* Configurations the kernel does not take are refused.
"""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff
from tests.test_gpu_batch import initial_states, ragged_systems, split, system_params, vicsek_params

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def pair(p, systems, x, y, th):
    a = S.Engine(p.copy().update(kernel_variant=S.KERNEL_PERSIST), systems)
    b = S.Engine(p.copy().update(kernel_variant=S.KERNEL_SLOT), systems)
    a.set_state(x, y, th)
    b.set_state(x, y, th)
    return a, b


def assert_same_state(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in ("xw", "yw", "xu", "yu", "theta"):
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("integrator", [S.RK2, S.RK4, S.EULER])
def test_period_batch_bitwise_equal_to_multilaunch(integrator):
    p = vicsek_params(integrator=integrator, rebuild=S.REBUILD_PERIOD, rebuild_period=9)
    systems = ragged_systems(7, seed=30 + integrator, n_lo=300, n_hi=4096)
    x, y, th = initial_states(systems, 50)
    a, b = pair(p, systems, x, y, th)
    for chunk in (40, 1, 59):  # several launches, one of a single step
        a.multi_step(chunk)
        b.multi_step(chunk)
        assert_same_state(a, b)
    assert a.info().steps == b.info().steps == 100
    assert a.info().t == b.info().t


def test_period_single_box_and_repulsion_bitwise():
    p = S.default_params(S.MODEL_REPULSIVE)
    p.update(n=3000, Lx=25.0, Ly=25.0, dt=0.005, v0=1.0, R_align=1.0, repulsion=S.REP_HARMONIC, A_rep=10.0,
             sigma_rep=1.0, rc_rep=1.0, D_rot=0.2, integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_PERIOD,
             rebuild_period=6, seed=4)
    x, y, th = S.random_state(3, p.n, p.Lx, p.Ly)
    a, b = pair(p, None, x, y, th)
    a.multi_step(150)
    b.multi_step(150)
    assert_same_state(a, b)


def test_capacity_overflow_inside_the_kernel():
    """Dense systems with the smallest list capacity: the kernel stops the overflowing systems,
    the lists grow and the launch is repeated; the result is that of the multi-launch engine."""
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=5, ell_capacity=12)
    systems = [S.system(n, 0.5 * math.sqrt(n / 4.0), v0=0.5, g=1.0, D_rot=0.2, seed=9 + n) for n in (500, 2000)]
    x, y, th = initial_states(systems, 3)
    a, b = pair(p, systems, x, y, th)
    a.multi_step(60)
    b.multi_step(60)
    assert_same_state(a, b)
    assert a.info().ell_capacity > 12


def test_samples_recorded_in_kernel_equal_multilaunch():
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=7, D_rot=0.3)
    systems = ragged_systems(4, seed=8, n_hi=3000)
    x, y, th = initial_states(systems, 2)
    a, b = pair(p, systems, x, y, th)
    ga = a.multi_step_sampled(280, 20)
    gb = b.multi_step_sampled(280, 20)
    assert np.array_equal(ga, gb)
    assert np.array_equal(a.batch_polarization(), ga[-1])


@pytest.mark.parametrize("integrator", [S.RK2, S.EULER])
def test_displacement_batch_matches_oracles(oracle_mod, integrator):
    p = vicsek_params(integrator=integrator)
    systems = ragged_systems(5, seed=60 + integrator, n_hi=4096)
    x, y, th = initial_states(systems, 70)
    a, b = pair(p, systems, x, y, th)
    a.multi_step(100)
    b.multi_step(100)
    got, ref = split(a, a.get_state()), split(b, b.get_state())
    for k, s in enumerate(systems):
        o = oracle_mod.Oracle(system_params(p, s))
        lo, hi = a.offsets[k], a.offsets[k + 1]
        o.set_state(x[lo:hi], y[lo:hi], th[lo:hi])
        o.multi_step(100)
        w = o.get_state()
        err = max(np.abs(got[k]["xu"] - w["xu"]).max(), np.abs(got[k]["yu"] - w["yu"]).max(),
                  angle_diff(got[k]["theta"], w["theta"]).max())
        assert err < TRAJ_TOL, (k, err)
        assert np.abs(got[k]["xu"] - ref[k]["xu"]).max() < TRAJ_TOL
    assert a.info().rebuilds > len(systems)  # displacement rebuilds happened inside the kernel


def test_refused_configurations():
    p = vicsek_params(kernel_variant=S.KERNEL_PERSIST)
    with pytest.raises(S.SppError):
        S.Engine(p, [S.system(5000, 40.0, v0=0.5, g=1.0, D_rot=0.1, seed=1)])  # above 4096 particles
    with pytest.raises(S.SppError):
        S.Engine(p.copy().update(noise_mode=S.NOISE_INJECTED, n=500, Lx=10.0, Ly=10.0))
    q = S.default_params(S.MODEL_ABP).update(n=200, Lx=10.0, Ly=10.0, kernel_variant=S.KERNEL_PERSIST)
    with pytest.raises(S.SppError):
        S.Engine(q)

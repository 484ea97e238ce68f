"""Pins of the oracle's active Brownian model and ring membrane by closed forms.

* A regular ring at the bond rest length is force free (bonds at rest; the bending energy depends
  on angles only, so a regular polygon is stationary by symmetry and scale invariance).
* Ring forces are minus the gradient of the ring energy (central differences).
* An overdamped square ring relaxes its side length as s - l = (s0 - l) exp(-4 m k sin^2(pi/4) t).
* A free swimmer without noise moves at speed mob_par * F0 along its heading; with an intrinsic
  torque its RK2 positions lie on a circle of radius h v / (2 sin(w h / 2)), w = mob_rot * torque0.
"""
import math

import numpy as np
import pytest

import spp_amd as S


def abp_params(**kw):
    p = S.default_params(S.MODEL_ABP)
    p.update(integrator=S.RK2, dt=1e-3, D_rot=0.0, repulsion=S.REP_NONE, r_list=1.3, k_bond=100.0,
             l_bond=1.0, kappa_bend=0.0, n_membrane=0, seed=5)
    return p.update(**kw)


def ring(n, radius, phase=0.0):
    a = phase + 2 * np.pi * np.arange(n) / n
    return radius * np.cos(a), radius * np.sin(a)


def test_regular_ring_at_rest_is_force_free(oracle_mod):
    n = 64
    R = 0.5 / math.sin(math.pi / n)  # side 1 = l_bond
    p = abp_params(n=n, n_membrane=n, Lx=R + 4, Ly=R + 4, kappa_bend=10.0)
    o = oracle_mod.Oracle(p)
    x, y = ring(n, R)
    o.set_state(x, y, np.zeros(n))
    o.interact()
    fx, fy, tau = o.get_forces()
    assert np.abs(fx).max() < 1e-11 and np.abs(fy).max() < 1e-11
    # stretched regular ring: bonds pull inwards, bending still contributes nothing
    p2 = abp_params(n=n, n_membrane=n, Lx=R + 4, Ly=R + 4, kappa_bend=0.0)
    o2 = oracle_mod.Oracle(p2)
    o3 = oracle_mod.Oracle(p2.copy().update(kappa_bend=10.0))
    x, y = ring(n, 1.1 * R)
    for q in (o2, o3):
        q.set_state(x, y, np.zeros(n))
        q.interact()
    a, b = o2.get_forces(), o3.get_forces()
    assert np.abs(a[0] - b[0]).max() < 1e-11 and np.abs(a[1] - b[1]).max() < 1e-11
    radial = (a[0] * x + a[1] * y) / np.hypot(x, y)
    want = -2 * 100.0 * (2 * 1.1 * R * math.sin(math.pi / n) - 1.0) * math.sin(math.pi / n)
    assert np.allclose(radial, want, rtol=1e-10)


@pytest.mark.parametrize("kappa", [0.0, 10.0])
def test_ring_forces_are_minus_energy_gradient(oracle_mod, kappa):
    n = 24
    R = 0.5 / math.sin(math.pi / n)
    rng = np.random.default_rng(7)
    x, y = ring(n, R)
    x = x + 0.2 * rng.standard_normal(n)
    y = y + 0.2 * rng.standard_normal(n)
    # place the ring across the periodic boundary to exercise the minimum image
    L = R + 3
    x = np.where(x > L - 0.5, x - 2 * L, x) + (L - 1.0)
    x = x - 2 * L * np.floor(x / (2 * L) + 0.5)
    p = abp_params(n=n, n_membrane=n, Lx=L, Ly=L, kappa_bend=kappa, k_bond=50.0, l_bond=0.9)
    o = oracle_mod.Oracle(p)
    o.set_state(x, y, np.zeros(n))
    o.interact()
    fx, fy, _ = o.get_forces()
    h = 1e-6
    gx, gy = np.empty(n), np.empty(n)
    for i in range(n):
        for arr, g in ((x, gx), (y, gy)):
            old = arr[i]
            arr[i] = old + h
            o.set_state(x, y, np.zeros(n))
            ep = o.ring_energy()
            arr[i] = old - h
            o.set_state(x, y, np.zeros(n))
            em = o.ring_energy()
            arr[i] = old
            g[i] = -(ep - em) / (2 * h)
    scale = np.abs(np.concatenate([gx, gy])).max()
    assert np.abs(fx - gx).max() < 1e-6 * scale and np.abs(fy - gy).max() < 1e-6 * scale


def test_square_ring_relaxation_closed_form(oracle_mod):
    n, s0, l0, K, m, dt, steps = 4, 1.3, 1.0, 100.0, 0.5, 1e-4, 200
    R0 = s0 / (2 * math.sin(math.pi / n))
    p = abp_params(n=n, n_membrane=n, Lx=4.0, Ly=4.0, k_bond=K, l_bond=l0, dt=dt, kappa_bend=10.0,
                   mob_par=m, mob_perp=m)
    o = oracle_mod.Oracle(p)
    x, y = ring(n, R0, phase=0.3)
    o.set_state(x, y, np.zeros(n))
    o.multi_step(steps)
    st = o.get_state()
    side = np.hypot(np.diff(np.append(st["xu"], st["xu"][0])), np.diff(np.append(st["yu"], st["yu"][0])))
    gamma = 4 * m * K * math.sin(math.pi / n) ** 2
    want = l0 + (s0 - l0) * math.exp(-gamma * steps * dt)
    assert np.allclose(side, want, rtol=1e-5)


@pytest.mark.parametrize("integrator", [S.EULER, S.RK2])
def test_free_swimmer_straight_line(oracle_mod, integrator):
    p = abp_params(n=1, Lx=5.0, Ly=5.0, v0=2.0, mob_par=0.7, mob_perp=0.2, integrator=integrator, dt=0.01)
    o = oracle_mod.Oracle(p)
    o.set_state(np.array([0.1]), np.array([-0.3]), np.array([0.4]))
    o.multi_step(50)
    st = o.get_state()
    v = 0.7 * 2.0
    assert abs(st["xu"][0] - (0.1 + v * 0.5 * math.cos(0.4))) < 1e-12
    assert abs(st["yu"][0] - (-0.3 + v * 0.5 * math.sin(0.4))) < 1e-12


def test_chiral_swimmer_circle(oracle_mod):
    h, F0, m, k, tq = 0.01, 2.0, 0.5, 1.0, 0.8
    p = abp_params(n=1, Lx=6.0, Ly=6.0, v0=F0, mob_par=m, mob_perp=m, mob_rot=k, torque0=tq, dt=h)
    o = oracle_mod.Oracle(p)
    o.set_state(np.array([0.0]), np.array([0.0]), np.array([0.0]))
    pts = []
    for _ in range(40):
        o.multi_step(5)
        st = o.get_state()
        pts.append((st["xu"][0], st["yu"][0]))
    pts = np.array(pts)
    w, v = k * tq, m * F0
    Rd = h * v / (2 * math.sin(w * h / 2))
    # RK2 steps are chords of length h v along the midpoint headings n w h + w h / 2: the points
    # are vertices of a regular polygon inscribed in a circle of radius Rd centred at (0, Rd)
    c = np.array([0.0, Rd])
    r = np.hypot(pts[:, 0] - c[0], pts[:, 1] - c[1])
    assert np.abs(r - Rd).max() < 1e-10

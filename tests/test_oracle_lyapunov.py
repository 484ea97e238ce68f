"""Pins of the CPU oracle's Lyapunov step (oracle/lyapunov.py) by textbook routines and closed forms.

* Gram-Schmidt in the reference's order equals Householder QR (numpy.linalg.qr) up to column signs,
  and the growth factors are |R_ii|.
* The difference and wrap formulas equal the C oracle's Periodic_Transform and the mathematical
  definition of the angle wrap; a deviation added and taken again round-trips.
* Two aligned particles at rest: the exponents of the full 6-dimensional tangent space are
  {0, 0, 0, 0, 0, -2 g / (pi R_a^2)} (the alignment torque's Jacobian), and their sum is the
  phase-space contraction rate exactly.
"""
import math

import numpy as np
import pytest

import spp_amd as S


def test_gram_schmidt_equals_householder_qr():
    import oracle.lyapunov as OL

    rng = np.random.default_rng(4)
    for K, M in [(1, 5), (3, 30), (6, 6), (8, 200)]:
        V = rng.standard_normal((K, M))
        U, growth = OL.gram_schmidt(V)
        Q, R = np.linalg.qr(V.T)
        sign = np.sign(np.diag(R))
        assert np.allclose(U, (Q * sign).T, atol=1e-10)
        assert np.allclose(growth, np.abs(np.diag(R)), rtol=1e-12)
        assert np.allclose(U @ U.T, np.eye(K), atol=1e-12)


def test_gram_schmidt_keeps_index_order():
    """Vector i only mixes vectors 0..i (R upper triangular): the first direction is v_0 itself."""
    import oracle.lyapunov as OL

    V = np.array([[3.0, 0, 0, 0], [1.0, 2.0, 0, 0], [5.0, 1.0, 7.0, 0]])
    U, growth = OL.gram_schmidt(V)
    assert np.allclose(U, np.eye(4)[:3])
    assert np.allclose(growth, [3.0, 2.0, 7.0])


def test_wraps_match_reference_formulas(oracle_mod):
    import oracle.lyapunov as OL

    rng = np.random.default_rng(1)
    L2 = 12.5
    x = np.concatenate([rng.uniform(-2 * L2, 2 * L2, 1000), [L2 / 2, -L2 / 2, 0.0, L2, -L2]])
    got = OL.ref_wrap(x, L2)
    want = np.array([oracle_mod.periodic(v, L2) for v in x])
    assert np.array_equal(got, want)
    t = np.concatenate([rng.uniform(-20, 20, 1000), [math.pi, -math.pi, 0.0, 3 * math.pi]])
    w = OL.wrap_theta_diff(t)
    assert np.all(w > -math.pi) and np.all(w <= math.pi)
    assert np.allclose(np.cos(w), np.cos(t), atol=1e-12) and np.allclose(np.sin(w), np.sin(t), atol=1e-12)
    assert OL.wrap_theta_diff(math.pi) == math.pi and OL.wrap_theta_diff(-math.pi) == math.pi


def test_deviation_round_trip_across_the_boundary():
    import oracle.lyapunov as OL

    rng = np.random.default_rng(2)
    n, L = 64, 5.0
    base = {"xw": rng.uniform(-L, L, n), "yw": rng.uniform(-L, L, n), "theta": rng.uniform(-3, 3, n)}
    base["xw"][:4] = [L - 1e-9, -L, L - 1e-7, -L + 1e-8]  # next to the edges
    base["xu"], base["yu"] = base["xw"] + 20, base["yw"] - 20
    d = 1e-6 * rng.standard_normal((3, n))
    clone = OL.add_deviation(base, d, L, L)
    assert np.all(np.abs(clone["xw"]) <= L)
    back = OL.deviation(base, clone, L, L)
    assert np.allclose(back, d, atol=1e-14)
    assert np.allclose(clone["xu"] - base["xu"], d[0], atol=1e-14)


def _pair_params(model_integrator=S.RK4):
    p = S.default_params(S.MODEL_REPULSIVE)
    return p.update(n=2, Lx=2.0, Ly=2.0, v0=0.0, g=1.0, R_align=1.0, repulsion=S.REP_NONE, D_rot=0.0,
                    integrator=model_integrator, dt=0.01, r_list=1.3, seed=3)


def _pair_state():
    return np.array([0.0, 0.5]), np.array([0.1, -0.2]), np.array([0.3, 0.3])


def test_aligned_pair_spectrum_closed_form(oracle_mod):
    """Full spectrum of two aligned particles at rest by the oracle: one direction contracts at
    2 g / (pi R_a^2), five are neutral."""
    import oracle.lyapunov as OL

    p = _pair_params()
    x, y, th = _pair_state()
    K, eps, m, intervals = 6, 1e-6, 10, 200
    systems = [oracle_mod.Oracle(p) for _ in range(K + 1)]
    for s in systems:
        s.set_state(x, y, th)
    lo = OL.LyapunovOracle(systems, p.Lx, p.Ly, eps)
    V0 = np.random.default_rng(3).uniform(-1, 1, (K, 3, 2))
    U, _ = OL.gram_schmidt(V0)
    b = systems[0].get_state()
    for k in range(K):
        c = OL.add_deviation(b, eps * U[k], p.Lx, p.Ly)
        systems[k + 1].set_state(c["xw"], c["yw"], c["theta"], c["xu"], c["yu"])
    half = None
    for it in range(intervals):
        lo.multi_step(m)
        lo.renormalize()
        if it == intervals // 2 - 1:
            half = lo.logsum.copy()
    T = intervals * m * p.dt
    rate = 2 * p.g / (math.pi * p.R_align ** 2)
    # the sum is the log-determinant of the tangent map: exact from the first interval on
    assert abs(lo.logsum.sum() / T + rate) < 1e-6 * rate
    # individual exponents once the directions have aligned (slope over the second half)
    slope = (lo.logsum - half) / (T / 2)
    assert abs(slope[-1] + rate) < 1e-4
    assert np.all(np.abs(slope[:-1]) < 1e-4)

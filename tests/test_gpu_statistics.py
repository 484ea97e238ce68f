"""GPU: steady-state statistics of the engine against the CPU oracle, within statistical error.

Identical-noise trajectories agree to 1e-9 only until chaos separates them; beyond that the
engine and the oracle must agree as statistical ensembles.  Each test runs independent systems
(different seeds and starting states) on both sides, reduces every system to one time average,
and compares the two ensemble means with Welch's t statistic, so the samples entering the error
are independent by construction.

* polarisation of the continuous-time Vicsek model (particle.h:268-338) in its ordered phase;
  doubling D (0.3 -> 0.6) moves the mean from about 0.67 to 0.30, over 20 standard errors;
* the pair distribution of the repulsive model with the harmonic law at noise D = 0.2 (its
  first-shell peak is set by the repulsion);
* the largest Lyapunov exponent of the repulsive model (RK4, Benettin renormalisation every 10
  steps; parallel/lyapunovbox.h:128-194), engine tangent vectors against the oracle's restatement.

Bounds: |t| < 5 for the polarisation (16 + 16 systems) and for lambda_1 (8 + 8 replicas), and
< 6 per bin of at least 20 counts for g(r) (8 + 8 systems, 13 bins); the same comparisons
between two independent oracle ensembles give |t| of 0.5-2.
"""
import math

import numpy as np
import pytest

import spp_amd as S

pytestmark = pytest.mark.gpu


def welch_t(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    se = np.sqrt(a.var(axis=0, ddof=1) / a.shape[0] + b.var(axis=0, ddof=1) / b.shape[0])
    return (a.mean(axis=0) - b.mean(axis=0)) / se


# ---------------------------------------------------------------------------------------------
# Polarisation: 16 continuous-Vicsek boxes of 256 particles, D = 0.3 (P about 0.67)
# ---------------------------------------------------------------------------------------------
M_POL, N_POL, L_POL, D_POL = 16, 256, 8.0, 0.3
BURN, SAMPLES, EVERY = 2000, 100, 20


def vicsek_params(seed):
    p = S.default_params(S.MODEL_CONTINUOUS)
    return p.update(n=N_POL, Lx=L_POL, Ly=L_POL, dt=0.01, v0=0.5, g=1.0, R_align=1.0, alpha=0.0,
                    D_rot=D_POL, integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT, seed=seed)


def aligned_start(oracle_mod, seed):
    x, y, th = oracle_mod.random_state(seed, N_POL, L_POL, L_POL)
    return x, y, 0.3 * (th / math.pi - 1.0)  # headings within 0.3 of zero: no ordering transient


def test_polarization_ensemble_matches_oracle(oracle_mod):
    # engine: one batched launch sequence over 16 independent boxes, sampled on the device
    systems = [S.system(N_POL, L_POL, v0=0.5, g=1.0, D_rot=D_POL, seed=500 + k) for k in range(M_POL)]
    starts = [aligned_start(oracle_mod, 500 + k) for k in range(M_POL)]
    e = S.Engine(vicsek_params(500), systems)
    e.set_state(*[np.concatenate([s[c] for s in starts]) for c in range(3)])
    e.multi_step(BURN)
    gpu = e.multi_step_sampled(SAMPLES * EVERY, EVERY).mean(axis=0)
    # oracle: 16 other boxes, one after another
    cpu = []
    for k in range(M_POL):
        o = oracle_mod.Oracle(vicsek_params(900 + k))
        o.set_state(*aligned_start(oracle_mod, 900 + k))
        o.multi_step(BURN)
        ps = []
        for _ in range(SAMPLES):
            o.multi_step(EVERY)
            ps.append(o.polarization()[0])
        cpu.append(np.mean(ps))
    t = welch_t(gpu, cpu)
    print("P engine %.4f, oracle %.4f, t %.2f" % (gpu.mean(), np.mean(cpu), t))
    assert 0.4 < gpu.mean() < 0.9  # the ordered phase
    assert abs(t) < 5


# ---------------------------------------------------------------------------------------------
# g(r): 8 repulsive boxes of 512 particles, harmonic k = 10 over radius 1, D = 0.2
# ---------------------------------------------------------------------------------------------
M_GR, N_GR, R_MAX, NBINS = 8, 512, 1.3, 13
L_GR = 0.5 * math.sqrt(N_GR)


def repulsive_params(seed):
    p = S.default_params(S.MODEL_REPULSIVE)
    return p.update(n=N_GR, Lx=L_GR, Ly=L_GR, dt=0.005, v0=1.0, g=1.0, R_align=1.0, repulsion=S.REP_HARMONIC,
                    A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, D_rot=0.2, integrator=S.RK2, r_list=1.3,
                    rebuild=S.REBUILD_DISPLACEMENT, seed=seed)


def oracle_counts(oracle_mod, x, y):
    """Ordered-pair counts in NBINS bins over [0, R_MAX), by brute force over minimum images."""
    i, j = oracle_mod.brute_pairs(x, y, L_GR, L_GR, R_MAX)
    dx, dy = x[i] - x[j], y[i] - y[j]
    dx -= 2 * L_GR * np.rint(dx / (2 * L_GR))
    dy -= 2 * L_GR * np.rint(dy / (2 * L_GR))
    b = np.minimum((np.sqrt(dx * dx + dy * dy) / (R_MAX / NBINS)).astype(np.int64), NBINS - 1)
    return 2.0 * np.bincount(b, minlength=NBINS)


def test_pair_distribution_ensemble_matches_oracle(oracle_mod):
    gpu, cpu = [], []
    for k in range(M_GR):
        x, y, th = oracle_mod.random_state(600 + k, N_GR, L_GR, L_GR)
        e = S.Engine(repulsive_params(600 + k))
        e.set_state(x, y, th)
        e.multi_step(1000)
        h = []
        for _ in range(10):
            e.multi_step(100)
            h.append(e.pair_distribution(R_MAX, NBINS)[2])
        gpu.append(np.mean(h, axis=0))

        x, y, th = oracle_mod.random_state(700 + k, N_GR, L_GR, L_GR)
        o = oracle_mod.Oracle(repulsive_params(700 + k))
        o.set_state(x, y, th)
        o.multi_step(1000)
        h = []
        for _ in range(10):
            o.multi_step(100)
            s = o.get_state()
            h.append(oracle_counts(oracle_mod, s["xw"], s["yw"]))
        cpu.append(np.mean(h, axis=0))
    gpu, cpu = np.array(gpu), np.array(cpu)
    full = (gpu.mean(axis=0) >= 20) & (cpu.mean(axis=0) >= 20)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = welch_t(gpu, cpu)
    print("counts engine", np.round(gpu.mean(axis=0), 1))
    print("counts oracle", np.round(cpu.mean(axis=0), 1))
    print("t", np.round(t, 2))
    assert full.sum() >= 6
    assert np.all(np.abs(t[full]) < 6)
    # the repulsion empties the core: on average under one pair per snapshot closer than 0.4
    assert gpu.mean(axis=0)[:4].sum() < 1 and cpu.mean(axis=0)[:4].sum() < 1


# ---------------------------------------------------------------------------------------------
# Largest Lyapunov exponent: 8 replicas of 256 repulsive particles, RK4, renormalised every 10
# steps (Benettin, amplitude 1e-6), over 2000 steps
# ---------------------------------------------------------------------------------------------
R_LY, N_LY, L_LY, D_LY, EPS_LY, STEPS_LY, EVERY_LY = 8, 256, 8.0, 0.05, 1e-6, 2000, 10


def lyapunov_params(seed):
    return S.default_params(S.MODEL_REPULSIVE).update(
        n=N_LY, Lx=L_LY, Ly=L_LY, v0=1.0, g=1.0, D_rot=D_LY, seed=seed, dt=0.005, R_align=1.0,
        repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, integrator=S.RK4, r_list=1.3,
        rebuild=S.REBUILD_DISPLACEMENT)


def test_largest_lyapunov_exponent_ensemble_matches_oracle(oracle_mod):
    import oracle.lyapunov as OL

    # engine: replicas x (base + one clone) in one batched engine
    systems = [S.system(N_LY, L_LY, v0=1.0, g=1.0, D_rot=D_LY, seed=500 + r) for r in range(R_LY) for _ in range(2)]
    xs, ys, ts = [], [], []
    for r in range(R_LY):
        x, y, t = oracle_mod.random_state(500 + r, N_LY, L_LY, L_LY)
        xs += [x, x]
        ys += [y, y]
        ts += [t, t]
    e = S.Engine(lyapunov_params(500), systems)
    e.set_state(np.concatenate(xs), np.concatenate(ys), np.concatenate(ts))
    e.lyapunov_init(R_LY, 1, EPS_LY, rescale=True, seed=4)
    e.lyapunov_run(STEPS_LY, EVERY_LY)
    sums, T = e.lyapunov_get()
    gpu = sums[:, 0] / T
    # oracle: other replicas, the clone started at a random deviation of length 1e-6
    cpu = []
    for r in range(R_LY):
        p = lyapunov_params(900 + r)
        x, y, th = oracle_mod.random_state(900 + r, N_LY, L_LY, L_LY)
        base = oracle_mod.Oracle(p)
        base.set_state(x, y, th)
        clone = oracle_mod.Oracle(p)
        d = np.random.default_rng(900 + r).standard_normal((3, N_LY))
        s = OL.add_deviation(base.get_state(), d * (EPS_LY / np.sqrt((d * d).sum())), L_LY, L_LY)
        clone.set_state(s["xw"], s["yw"], s["theta"], s["xu"], s["yu"])
        lo = OL.LyapunovOracle([base, clone], L_LY, L_LY, EPS_LY, True)
        for _ in range(STEPS_LY // EVERY_LY):
            lo.multi_step(EVERY_LY)
            lo.renormalize()
        cpu.append(lo.logsum[0] / (STEPS_LY * p.dt))
    t = welch_t(gpu, cpu)
    print("lambda_1 engine %.3f (sd %.3f), oracle %.3f (sd %.3f), t %.2f" % (gpu.mean(), gpu.std(ddof=1), np.mean(cpu),
                                                                          np.std(cpu, ddof=1), t))
    assert abs(T - STEPS_LY * 0.005) < 1e-9
    assert gpu.mean() > 0.3  # chaotic
    assert abs(t) < 5

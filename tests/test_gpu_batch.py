"""GPU: batched engines (many independent boxes in one engine) and Lyapunov clones.

Batches: every system of a batch must evolve exactly as it would alone -- against an independent
oracle run per system (100 steps within 1e-9), bitwise against a single-system engine with the same
rebuild steps, and bitwise between identical systems (the reference's zero-deviation direction 0
replays the base trajectory).

Lyapunov: accumulated log growth and the orthonormal directions against the oracle's restatement of
the reference's Gram-Schmidt (oracle/lyapunov.py) from identical clone states; orthonormality; the
closed-form spectrum of two aligned particles at rest.
"""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff

pytestmark = pytest.mark.gpu

TRAJ_TOL = 1e-9


def vicsek_params(**kw):
    """Continuous-time Vicsek (R = 1, v0 = 0.5, g = 1), RK2 dt = 0.01, list radius 1.3."""
    p = S.default_params(S.MODEL_CONTINUOUS)
    p.update(dt=0.01, R_align=1.0, alpha=0.0, integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT)
    return p.update(**kw)


def ragged_systems(count, seed, n_lo=300, n_hi=5000):
    """Ragged ensemble: N log-uniform, density U[0.25, 4], D U[0, 0.5], v0 0.5, distinct seeds."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        rho = rng.uniform(0.25, 4.0)
        L = 0.5 * math.sqrt(n / rho)
        out.append(S.system(n, L, v0=0.5, g=1.0, D_rot=rng.uniform(0, 0.5), seed=100 + k))
    return out


def system_params(p, s):
    return p.copy().update(n=s.n, Lx=s.Lx, Ly=s.Ly, v0=s.v0, g=s.g, D_rot=s.D_rot, seed=s.seed)


def initial_states(systems, seed):
    xs, ys, ts = [], [], []
    for k, s in enumerate(systems):
        x, y, t = S.random_state(seed + k, s.n, s.Lx, s.Ly)
        xs.append(x)
        ys.append(y)
        ts.append(t)
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ts)


def split(e, st):
    return [{k: v[e.offsets[i]:e.offsets[i + 1]] for k, v in st.items()} for i in range(len(e.sizes))]


# ---------------------------------------------------------------------------------------------
# Batches
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("integrator", [S.RK2, S.RK4, S.EULER])
def test_batch_matches_independent_oracles(oracle_mod, integrator):
    p = vicsek_params(integrator=integrator)
    systems = ragged_systems(6, seed=integrator)
    x, y, th = initial_states(systems, 40)
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    e.multi_step(100)
    got = split(e, e.get_state())
    pol = e.batch_polarization()
    for k, s in enumerate(systems):
        o = oracle_mod.Oracle(system_params(p, s))
        a, b = e.offsets[k], e.offsets[k + 1]
        o.set_state(x[a:b], y[a:b], th[a:b])
        o.multi_step(100)
        w = o.get_state()
        err = max(np.abs(got[k]["xu"] - w["xu"]).max(), np.abs(got[k]["yu"] - w["yu"]).max(),
                  angle_diff(got[k]["theta"], w["theta"]).max())
        assert err < TRAJ_TOL, (k, s.n, err)
        assert abs(pol[k] - o.polarization()[0]) < 1e-10


def test_batch_repulsive_matches_oracles(oracle_mod):
    """Repulsive Vicsek (harmonic k = 10 over rc = 1) boxes of different sizes and noise."""
    p = S.default_params(S.MODEL_REPULSIVE).update(
        dt=0.005, R_align=1.0, repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0,
        integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT)
    systems = [S.system(n, 0.5 * math.sqrt(n), v0=1.0, g=1.0, D_rot=D, seed=7 + k)
               for k, (n, D) in enumerate([(1024, 0.05), (4096, 0.2), (2500, 0.1)])]
    x, y, th = initial_states(systems, 3)
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    e.multi_step(100)
    got = split(e, e.get_state())
    for k, s in enumerate(systems):
        o = oracle_mod.Oracle(system_params(p, s))
        a, b = e.offsets[k], e.offsets[k + 1]
        o.set_state(x[a:b], y[a:b], th[a:b])
        o.multi_step(100)
        w = o.get_state()
        assert np.abs(got[k]["xu"] - w["xu"]).max() < TRAJ_TOL
        assert angle_diff(got[k]["theta"], w["theta"]).max() < TRAJ_TOL


def test_identical_systems_stay_bitwise_identical():
    """Same state and seed in three slots of a batch: the trajectories are bit-for-bit equal (the
    reference's zero-deviation direction replays the base)."""
    p = vicsek_params(D_rot=0.2)
    s = S.system(3000, 20.0, v0=0.5, g=1.0, D_rot=0.2, seed=9)
    x, y, th = S.random_state(5, 3000, 20.0, 20.0)
    e = S.Engine(p, [s, s, s])
    e.set_state(np.tile(x, 3), np.tile(y, 3), np.tile(th, 3))
    e.multi_step(300)
    st = split(e, e.get_state())
    for k in ("xw", "yw", "xu", "yu", "theta"):
        assert np.array_equal(st[0][k], st[1][k]) and np.array_equal(st[0][k], st[2][k])


def test_batch_system_equals_single_engine_bitwise():
    """With the same rebuild steps, a system inside a batch computes exactly what a single-system
    engine computes (same lists, same summation order, same noise)."""
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=8)
    systems = ragged_systems(4, seed=11)
    x, y, th = initial_states(systems, 70)
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    e.multi_step(64)
    got = split(e, e.get_state())
    for k, s in enumerate(systems):
        a, b = e.offsets[k], e.offsets[k + 1]
        one = S.Engine(system_params(p, s))
        one.set_state(x[a:b], y[a:b], th[a:b])
        one.multi_step(64)
        w = one.get_state()
        for key in ("xw", "yw", "xu", "yu", "theta"):
            assert np.array_equal(got[k][key], w[key]), (k, key)


@pytest.mark.parametrize("law,integrator", [(S.REP_HARMONIC, S.RK2), (S.REP_HARMONIC, S.RK4),
                                             (S.REP_HARMONIC, S.EULER), (S.REP_R12_TRUNCATED, S.RK2)])
def test_one_cutoff_kernels_equal_the_general_pair_path_bitwise(law, integrator):
    """Repulsion and alignment ending at the same radius (rc = R_a = 1, the benchmark's harmonic
    workload): a single repulsive engine then runs the stage kernels without the two per-pair cutoff
    tests, batched engines keep the general pair path.  Same rebuild steps, so the same lists and
    summation order: the single engine must reproduce its system inside a batch bit for bit."""
    p = S.default_params(S.MODEL_REPULSIVE).update(
        dt=0.005, R_align=1.0, repulsion=law, A_rep=10.0 if law == S.REP_HARMONIC else 2.0, sigma_rep=1.0,
        rc_rep=1.0, integrator=integrator, r_list=1.3, rebuild=S.REBUILD_PERIOD, rebuild_period=5)
    systems = [S.system(n, 0.5 * math.sqrt(n), v0=1.0, g=1.0, D_rot=D, seed=30 + k)
               for k, (n, D) in enumerate([(3000, 0.05), (1800, 0.2)])]
    x, y, th = initial_states(systems, 90)
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    e.multi_step(60)
    got = split(e, e.get_state())
    for k, s in enumerate(systems):
        a, b = e.offsets[k], e.offsets[k + 1]
        one = S.Engine(system_params(p, s))
        one.set_state(x[a:b], y[a:b], th[a:b])
        one.multi_step(60)
        w = one.get_state()
        for key in ("xw", "yw", "xu", "yu", "theta"):
            assert np.array_equal(got[k][key], w[key]), (k, key)


def c4a_ensemble(seed=4):
    """The ensemble workload: 256 Vicsek boxes, N log-uniform in [1024, 65536], density U[0.25, 4],
    D U[0, 0.5], v0 0.5 (about 3.9 million particles), uniform random starts."""
    rng = np.random.default_rng(seed)
    systems, xs = [], []
    for k in range(256):
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(65536)))))
        rho = rng.uniform(0.25, 4.0)
        L = 0.5 * math.sqrt(n / rho)
        systems.append(S.system(n, L, v0=0.5, g=1.0, D_rot=rng.uniform(0, 0.5), seed=100 + k))
        xs.append(S.random_state(200 + k, n, L, L))
    return systems, [np.concatenate([v[i] for v in xs]) for i in range(3)]


def test_c4a_ensemble_256_ragged_boxes(oracle_mod):
    """The full 256-box ensemble in one engine, run as two 100-step calls (RK2, dt 0.01,
    displacement rebuilds): every polarisation lies in [0, 1], each box's torques sum to zero
    (Newton-3), and the five smallest boxes follow independent oracle runs within 1e-9."""
    p = vicsek_params()
    systems, (x, y, th) = c4a_ensemble()
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    assert e.info().ell_capacity >= e.info().max_neighbors
    for _ in range(2):
        e.multi_step(100)
    pol = e.batch_polarization()
    assert pol.shape == (256,) and np.all(pol >= 0) and np.all(pol <= 1 + 1e-12)
    got = split(e, e.get_state())
    e.interact()
    tau = e.get_forces()[2]
    for k in range(256):
        a, b = e.offsets[k], e.offsets[k + 1]
        assert abs(tau[a:b].sum()) < 1e-9 * (b - a), k
    smallest = np.argsort([s.n for s in systems])[:5]
    for k in smallest:
        s = systems[k]
        a, b = e.offsets[k], e.offsets[k + 1]
        o = oracle_mod.Oracle(system_params(p, s))
        o.set_state(x[a:b], y[a:b], th[a:b])
        o.multi_step(200)
        w = o.get_state()
        err = max(np.abs(got[k]["xu"] - w["xu"]).max(), np.abs(got[k]["yu"] - w["yu"]).max(),
                  angle_diff(got[k]["theta"], w["theta"]).max())
        assert err < TRAJ_TOL, (k, s.n, err)
        assert abs(pol[k] - o.polarization()[0]) < 1e-10


def test_c4a_ensemble_systems_equal_single_engines_bitwise():
    """With fixed-period rebuilds every box of the 256-box ensemble has its rebuilds at the same
    steps as a single-box engine, so selected boxes (smallest, largest, densest, noisiest) are bit
    for bit what they compute alone."""
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=10)
    systems, (x, y, th) = c4a_ensemble()
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    e.multi_step(100)
    got = split(e, e.get_state())
    sizes = np.array([s.n for s in systems])
    dens = np.array([s.n / (4 * s.Lx * s.Ly) for s in systems])
    noise = np.array([s.D_rot for s in systems])
    for k in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmax(dens)), int(np.argmax(noise))}):
        s = systems[k]
        a, b = e.offsets[k], e.offsets[k + 1]
        one = S.Engine(system_params(p, s))
        one.set_state(x[a:b], y[a:b], th[a:b])
        one.multi_step(100)
        w = one.get_state()
        for key in ("xw", "yw", "xu", "yu", "theta"):
            assert np.array_equal(got[k][key], w[key]), (k, key)


def test_batch_neighbors_are_per_system():
    """No list entry crosses systems, and each system's list is its own brute-force pair set."""
    import oracle
    from tests.helpers import csr_pair_set, pair_set

    p = vicsek_params()
    systems = ragged_systems(5, seed=2)
    x, y, th = initial_states(systems, 1)
    e = S.Engine(p, systems)
    e.set_state(x, y, th)
    off, idx = e.get_neighbors()
    got = csr_pair_set(off, idx)
    want = set()
    for k, s in enumerate(systems):
        a, b = int(e.offsets[k]), int(e.offsets[k + 1])
        pi, pj = oracle.brute_pairs(x[a:b], y[a:b], s.Lx, s.Ly, p.r_list)
        want |= pair_set(pi + a, pj + a)
    assert got == want


def test_batch_rejects_bad_input():
    p = vicsek_params()
    with pytest.raises(S.SppError):
        S.Engine(p, [])
    with pytest.raises(S.SppError):  # box too small for three cells of r_list
        S.Engine(p, [S.system(10, 1.0)])
    with pytest.raises(S.SppError):  # a retired kernel variant
        S.Engine(p.copy().update(kernel_variant=2), [S.system(100, 5.0)])


# ---------------------------------------------------------------------------------------------
# Lyapunov clones
# ---------------------------------------------------------------------------------------------

def lyap_params(**kw):
    p = S.default_params(S.MODEL_REPULSIVE).update(
        dt=0.005, R_align=1.0, repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0,
        integrator=S.RK4, r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT)
    return p.update(**kw)


def lyap_engine(p, R, K, n, L, D, eps, rescale=True):
    systems = [S.system(n, L, v0=1.0, g=1.0, D_rot=D, seed=20 + r) for r in range(R) for _ in range(K + 1)]
    xs, ys, ts = [], [], []
    for r in range(R):
        x, y, t = S.random_state(50 + r, n, L, L)
        for _ in range(K + 1):
            xs.append(x)
            ys.append(y)
            ts.append(t)
    e = S.Engine(p, systems)
    e.set_state(np.concatenate(xs), np.concatenate(ys), np.concatenate(ts))
    e.lyapunov_init(R, K, eps, rescale=rescale, seed=4)
    return e, systems


def test_lyapunov_deep_variant_bitwise_equals_slot():
    """Twenty-five 16,384-particle systems (5 replicas x 5) exceed 393,216 particles, where the
    slot kernel keeps four neighbour records in flight and the deep variant eight; both add the
    pairs in list order, so states and growth sums are bit-for-bit equal."""
    R, K, n, L = 5, 4, 16384, 64.0
    lx, ly = S.triangle_lattice(n, L, L)
    out = []
    for v in (S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP):
        systems = [S.system(n, L, v0=1.0, g=1.0, D_rot=0.05, seed=20 + r) for r in range(R) for _ in range(K + 1)]
        ts = [S.random_state(50 + r, n, L, L)[2] for r in range(R) for _ in range(K + 1)]
        e = S.Engine(lyap_params(kernel_variant=v), systems)
        e.set_state(np.tile(lx, R * (K + 1)), np.tile(ly, R * (K + 1)), np.concatenate(ts))
        e.lyapunov_init(R, K, 1e-7, rescale=True, seed=4)
        e.lyapunov_run(40, 10)
        out.append((e.get_state(), e.lyapunov_get()[0]))
    for k in ("xw", "yw", "xu", "yu", "theta"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


def oracle_replicas(oracle_mod, e, p, systems, R, K, eps, rescale):
    import oracle.lyapunov as OL

    st = split(e, e.get_state())
    reps = []
    for r in range(R):
        group = []
        for c in range(K + 1):
            k = r * (K + 1) + c
            o = oracle_mod.Oracle(system_params(p, systems[k]))
            s = st[k]
            o.set_state(s["xw"], s["yw"], s["theta"], s["xu"], s["yu"])
            group.append(o)
        reps.append(OL.LyapunovOracle(group, systems[0].Lx, systems[0].Ly, eps, rescale))
    return reps


@pytest.mark.parametrize("rescale", [True, False])
def test_lyapunov_matches_oracle(oracle_mod, rescale):
    R, K, n, L, eps, m = 2, 3, 400, 10.0, 1e-6, 10
    p = lyap_params()
    e, systems = lyap_engine(p, R, K, n, L, 0.05, eps, rescale)
    reps = oracle_replicas(oracle_mod, e, p, systems, R, K, eps, rescale)
    for it in range(5):
        e.multi_step(m)
        g = e.lyapunov_renormalize(want=True)
        for r, lo in enumerate(reps):
            lo.multi_step(m)
            want = lo.renormalize()
            assert np.abs(g[r] - want).max() < 1e-6, (it, r, g[r], want)
    sums, T = e.lyapunov_get()
    assert abs(T - 5 * m * p.dt) < 1e-12
    for r, lo in enumerate(reps):
        Q = e.lyapunov_vectors(r)
        assert np.abs(Q - lo.U).max() < 1e-5


def test_lyapunov_directions_orthonormal():
    R, K = 3, 4
    e, _ = lyap_engine(lyap_params(), R, K, 2048, 0.5 * math.sqrt(2048), 0.1, 1e-7)
    for r in range(R):
        Q = e.lyapunov_vectors(r).reshape(K, -1)
        assert np.abs(Q @ Q.T - np.eye(K)).max() < 1e-13
    e.lyapunov_run(40, 10)
    for r in range(R):
        Q = e.lyapunov_vectors(r).reshape(K, -1)
        assert np.abs(Q @ Q.T - np.eye(K)).max() < 1e-13
    sums, T = e.lyapunov_get()
    assert np.all(np.isfinite(sums)) and abs(T - 40 * 0.005) < 1e-12


def test_lyapunov_base_trajectory_untouched():
    """Renormalising moves only the clones: each replica's base follows the trajectory of a plain
    batch with the same systems (fixed rebuild steps, so bitwise)."""
    R, K, n, L = 2, 2, 600, 12.0
    p = lyap_params(rebuild=S.REBUILD_PERIOD, rebuild_period=5)
    e, systems = lyap_engine(p, R, K, n, L, 0.1, 1e-6)
    plain = S.Engine(p, systems)
    xs = [S.random_state(50 + r, n, L, L) for r in range(R)]
    cat = lambda i: np.concatenate([xs[r][i] for r in range(R) for _ in range(K + 1)])
    plain.set_state(cat(0), cat(1), cat(2))
    for _ in range(3):
        e.multi_step(10)
        e.lyapunov_renormalize()
        plain.multi_step(10)
    a, b = split(e, e.get_state()), split(plain, plain.get_state())
    for r in range(R):
        k = r * (K + 1)
        for key in ("xu", "yu", "theta"):
            assert np.array_equal(a[k][key], b[k][key])


def test_lyapunov_aligned_pair_closed_form():
    """Two aligned particles at rest, full 6-vector spectrum: sum = -2 g/(pi R_a^2) (log-det of the
    tangent map) and, once aligned, one direction contracts at that rate, five are neutral."""
    p = S.default_params(S.MODEL_REPULSIVE).update(
        v0=0.0, g=1.0, R_align=1.0, repulsion=S.REP_NONE, D_rot=0.0, integrator=S.RK4, dt=0.01, r_list=1.3)
    K, eps, m, intervals = 6, 1e-6, 10, 200
    systems = [S.system(2, 2.0, v0=0.0, g=1.0, D_rot=0.0, seed=3) for _ in range(K + 1)]
    e = S.Engine(p, systems)
    x, y, th = np.array([0.0, 0.5]), np.array([0.1, -0.2]), np.array([0.3, 0.3])
    e.set_state(np.tile(x, K + 1), np.tile(y, K + 1), np.tile(th, K + 1))
    e.lyapunov_init(1, K, eps)
    e.lyapunov_run(intervals // 2 * m, m)
    half, _ = e.lyapunov_get()
    e.lyapunov_run(intervals // 2 * m, m)
    full, T = e.lyapunov_get()
    rate = 2 / math.pi
    assert abs(full.sum() / T + rate) < 1e-6 * rate
    slope = (full[0] - half[0]) / (T / 2)
    assert abs(slope[-1] + rate) < 1e-4
    assert np.all(np.abs(slope[:-1]) < 1e-4)


def test_lyapunov_rejects_inconsistent_groups():
    p = lyap_params()
    systems = [S.system(500, 11.0, seed=1), S.system(500, 11.0, seed=2)]
    e = S.Engine(p, systems)
    x, y, t = S.random_state(1, 1000, 11.0, 11.0)
    e.set_state(x, y, t)
    with pytest.raises(S.SppError):
        e.lyapunov_init(1, 1, 1e-6)  # clone has another noise seed
    with pytest.raises(S.SppError):
        e.lyapunov_init(2, 1, 1e-6)  # 2 x (1 + 1) != 2 systems


@pytest.mark.parametrize("batched", [False, True])
def test_sampled_steps_equal_stepwise_polarization(batched):
    """Samples recorded on the device during one call equal polarisations read after each chunk of
    a step-by-step run (fixed rebuild period: identical rebuild steps, so identical states)."""
    p = vicsek_params(rebuild=S.REBUILD_PERIOD, rebuild_period=7, D_rot=0.3)
    systems = ragged_systems(3, seed=4) if batched else None
    if batched:
        x, y, th = initial_states(systems, 9)
    else:
        p.update(n=2000, Lx=20.0, Ly=20.0, seed=5)
        x, y, th = S.random_state(9, 2000, 20.0, 20.0)
    a, b = S.Engine(p, systems), S.Engine(p, systems)
    a.set_state(x, y, th)
    b.set_state(x, y, th)
    got = a.multi_step_sampled(300, 20)
    want = []
    for _ in range(15):
        b.multi_step(20)
        want.append(b.batch_polarization() if batched else [b.polarization()[0]])
    assert got.shape == (15, len(a.sizes))
    if batched:  # the same per-system reduction
        assert np.array_equal(got, np.array(want))
    else:  # one workgroup's sum against the block-partial sum of spp_polarization
        assert np.abs(got - np.array(want)).max() < 1e-14

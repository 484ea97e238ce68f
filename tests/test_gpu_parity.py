"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bars (from the path's specification): neighbour sets equal exactly; single-step forces and
torques within 1e-12 norm-wise relative; 100-step trajectories under the identical noise stream
within 1e-9; statistics within their error bars.
"""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff, c1_params, c2_params, csr_pair_set, pair_set, ref_params, rel_err, vec_rel_err

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9


def _pair(p, x, y, th, oracle_mod):
    e = S.Engine(p)
    e.set_state(x, y, th)
    o = oracle_mod.Oracle(p)
    o.set_state(x, y, th)
    return e, o


def _lattice(n_side, L, seed, jitter=0.05):
    import oracle

    g = (np.arange(n_side) + 0.5) * (2 * L / n_side) - L
    x, y = [a.ravel() for a in np.meshgrid(g, g)]
    jx, jy, th = oracle.random_state(seed, n_side * n_side, jitter, jitter)
    return x + jx, y + jy, th


# ---------------------------------------------------------------------------------------------
# Neighbour sets
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n,L,rl", [(4096, 32.0, 1.3), (20000, 40.0, 1.3), (500, 2.0, 1.3), (3000, 8.0, 2.6)])
def test_neighbor_set_exact_mode(oracle_mod, n, L, rl):
    p = c2_params(n=n, L=L, r_list=rl)
    x, y, th = oracle_mod.random_state(31, n, L, L)
    e, o = _pair(p, x, y, th, oracle_mod)
    off, idx = e.get_neighbors()
    got = csr_pair_set(off, idx)
    assert got == pair_set(*o.half_list())
    assert idx.size == 2 * len(got)  # full list is symmetric, no duplicates
    pi, pj = oracle_mod.brute_pairs(x, y, L, L, rl)
    assert got == pair_set(pi, pj)


def test_neighbor_set_reference_quirk_mode(oracle_mod):
    """Reference grid and binning (113 cells over side 128, rv = 1.0625, serial/box.h:116-117):
    the GPU list equals the reference's Verlet list, including the pairs the quirk drops."""
    L, n, rv = 64.0, 16000, 1.0625
    p = c2_params(n=n, L=L, r_list=rv, ncell_x=113, ncell_y=113, binning=S.BIN_REF_QUIRK,
                  R_align=1.0, rc_rep=1.0)
    x, y, th = oracle_mod.random_state(9, n, L, L)
    e, o = _pair(p, x, y, th, oracle_mod)
    off, idx = e.get_neighbors()
    assert csr_pair_set(off, idx) == pair_set(*o.half_list())
    cx, cy = e.get_cells()
    ox, oy = o.get_cells()
    assert np.array_equal(cx, ox) and np.array_equal(cy, oy)


def test_neighbor_set_full_size_c2(oracle_mod):
    p = c2_params()
    x, y, th = oracle_mod.random_state(2, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    off, idx = e.get_neighbors()
    assert csr_pair_set(off, idx) == pair_set(*o.half_list())
    inf = e.info()
    assert inf.max_neighbors <= inf.ell_capacity
    assert abs(inf.mean_neighbors - math.pi * 1.3 ** 2) < 0.1


# ---------------------------------------------------------------------------------------------
# Single-step forces
# ---------------------------------------------------------------------------------------------

FORCE_CASES = {
    "c2_harmonic": lambda: c2_params(n=20000, L=64.0),
    "c1_continuous": lambda: c1_params(),
    "continuous_alpha": lambda: c1_params(alpha=0.5, g=3.0),
    "r12_truncated": lambda: c2_params(n=20000, L=64.0, repulsion=S.REP_R12_TRUNCATED, A_rep=2.0, rc_rep=1.05, R_align=1.1, r_list=1.4),
    "r12": lambda: c2_params(n=20000, L=64.0, repulsion=S.REP_R12, A_rep=0.5, rc_rep=1.0, sigma_rep=0.8),
    "yukawa_truncated": lambda: c2_params(n=20000, L=64.0, repulsion=S.REP_YUKAWA_TRUNCATED, A_rep=1.0, rc_rep=1.2, sigma_rep=0.5, r_list=1.5),
    "yukawa": lambda: c2_params(n=20000, L=64.0, repulsion=S.REP_YUKAWA, A_rep=1.0, rc_rep=1.0, sigma_rep=0.5),
    "no_repulsion": lambda: c2_params(n=20000, L=64.0, repulsion=S.REP_NONE),
}


@pytest.mark.parametrize("case", sorted(FORCE_CASES))
def test_single_interact_matches_oracle(oracle_mod, case):
    p = FORCE_CASES[case]()
    x, y, th = oracle_mod.random_state(17, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gf = e.get_forces()
    of = o.get_forces()
    fx_err = vec_rel_err(gf[0], gf[1], of[0], of[1])
    assert fx_err < FORCE_TOL, fx_err
    assert rel_err(gf[2], of[2]) < FORCE_TOL


def test_single_interact_full_size_c2(oracle_mod):
    p = c2_params()
    x, y, th = oracle_mod.random_state(5, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gf, of = e.get_forces(), o.get_forces()
    assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
    assert rel_err(gf[2], of[2]) < FORCE_TOL
    # Newton-3 invariants of the GPU result itself
    assert abs((gf[0] - np.cos(th)).sum()) < 1e-8
    assert abs(gf[2].sum()) < 1e-9


def _wrap(t):
    return t - 2 * np.pi * np.floor((t + np.pi) / (2 * np.pi))


# Ordered flocks (low noise) hold nearly equal headings: the torque's sin(theta_j - theta_i) is then
# tiny and must keep its *relative* accuracy (particle.h:750 takes the sine of the angle difference).
ALIGNED_HEADINGS = {
    "spread_1e-4": lambda rng, n: 0.3 + rng.uniform(-1e-4, 1e-4, n),
    "spread_1e-6": lambda rng, n: 0.3 + rng.uniform(-1e-6, 1e-6, n),
    "spread_1e-8": lambda rng, n: 0.3 + rng.uniform(-1e-8, 1e-8, n),
    # the same heading seen as +pi and -pi: differences near +-2 pi
    "across_wrap_1e-6": lambda rng, n: _wrap(np.pi + rng.uniform(-1e-6, 1e-6, n)),
    # two counter-moving lanes: differences near 0 and near +-pi
    "anti_aligned_1e-6": lambda rng, n: _wrap(0.3 + np.pi * rng.integers(0, 2, n) + rng.uniform(-1e-6, 1e-6, n)),
}


@pytest.mark.parametrize("model", ["c2_harmonic", "c1_continuous"])
@pytest.mark.parametrize("case", sorted(ALIGNED_HEADINGS))
def test_interact_near_aligned_headings(oracle_mod, case, model):
    p = c2_params(n=20000, L=64.0) if model == "c2_harmonic" else c1_params(n=20000, Lx=64.0, Ly=64.0)
    x, y, _ = oracle_mod.random_state(23, p.n, p.Lx, p.Ly)
    th = ALIGNED_HEADINGS[case](np.random.default_rng(7), p.n)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gf, of = e.get_forces(), o.get_forces()
    assert np.linalg.norm(of[2]) > 0
    err = rel_err(gf[2], of[2])
    assert err < FORCE_TOL, err
    assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL


def test_interact_fully_aligned_box_is_torque_free(oracle_mod):
    """Equal headings (a D = 0 flock): every torque is exactly zero on both sides, and 100 noiseless
    steps keep every heading bit-for-bit while the positions follow the oracle."""
    p = c2_params(n=20000, L=64.0, D=0.0)
    x, y, _ = oracle_mod.random_state(29, p.n, p.Lx, p.Ly)
    th = np.full(p.n, 0.3)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    assert np.all(e.get_forces()[2] == 0.0) and np.all(o.get_forces()[2] == 0.0)
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    assert np.all(gs["theta"] == 0.3)
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL and np.abs(gs["yu"] - os_["yu"]).max() < TRAJ_TOL


@pytest.mark.parametrize("dth", [1e-3, 1e-9, 1e-14, np.pi - 1e-9])
def test_two_particle_torque_closed_form(oracle_mod, dth):
    """Two particles inside the alignment radius: tau_1 = g sin(th_2 - th_1) / (pi R^2) exactly as
    the reference evaluates it (the difference rounded once), to a few ulps relative."""
    p = c2_params(n=2, L=8.0, repulsion=S.REP_NONE, g=1.0, R_align=1.0)
    th1 = 0.3
    th = np.array([th1, _wrap(th1 + dth)])
    e = S.Engine(p)
    e.set_state(np.array([0.1, 0.6]), np.array([0.2, 0.2]), th)
    e.interact()
    tau = e.get_forces()[2]
    want = math.sin(th[1] - th[0]) / math.pi
    assert abs(tau[0] - want) <= 4e-16 * abs(want)
    assert tau[1] == -tau[0]


# ---------------------------------------------------------------------------------------------
# Stages and steps with injected noise
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("integ", [S.EULER, S.RK2, S.RK4])
def test_split_stages_match_oracle(oracle_mod, integ):
    """Box::Interact + Move_* one call at a time (the drop-in granularity)."""
    p = c2_params(n=8192, L=48.0, integrator=integ, noise_mode=S.NOISE_INJECTED)
    x, y, th = oracle_mod.random_state(3, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    rng = np.random.default_rng(0)
    stages = [0] if integ == S.EULER else list(range(1, integ + 1))
    for _ in range(3):
        dn = rng.normal(0, 0.05, p.n)
        e.inject_noise(dn)
        o.set_noise(dn)
        for st in stages:
            e.interact()
            o.interact()
            gf, of = e.get_forces(), o.get_forces()
            assert rel_err(gf[2], of[2]) < FORCE_TOL
            assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
            e.move_stage(st)
            o.move_stage(st)
            gs, os_ = e.get_state(), o.get_state()
            assert np.abs(gs["xw"] - os_["xw"]).max() < 1e-12
            assert np.abs(gs["xu"] - os_["xu"]).max() < 1e-12
            assert angle_diff(gs["theta"], os_["theta"]).max() < 1e-12


@pytest.mark.parametrize("integ", [S.EULER, S.RK2, S.RK4])
def test_fused_step_matches_oracle(oracle_mod, integ):
    p = c2_params(n=8192, L=48.0, integrator=integ, noise_mode=S.NOISE_INJECTED)
    x, y, th = oracle_mod.random_state(4, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    rng = np.random.default_rng(1)
    for _ in range(5):
        dn = rng.normal(0, 0.05, p.n)
        e.inject_noise(dn)
        o.set_noise(dn)
        e.one_step()
        o.multi_step(1)  # with the oracle's own rebuild check
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < 1e-11
    assert np.abs(gs["yu"] - os_["yu"]).max() < 1e-11
    assert angle_diff(gs["theta"], os_["theta"]).max() < 1e-11


# ---------------------------------------------------------------------------------------------
# 100-step trajectories under the identical noise stream
# ---------------------------------------------------------------------------------------------

TRAJ_CASES = {
    "c1_vicsek_rk2": lambda: c1_params(),
    "c2_small_rk2": lambda: c2_params(n=16384, L=64.0),
    "c2_small_rk4": lambda: c2_params(n=16384, L=64.0, integrator=S.RK4),
    "c2_period_rebuild": lambda: c2_params(n=16384, L=64.0, rebuild=S.REBUILD_PERIOD, rebuild_period=10),
}


@pytest.mark.parametrize("D", [0.05, 0.2])
def test_trajectory_100_steps_full_size_c2(oracle_mod, D):
    """The headline workload itself: N = 262,144 in a box of side 512, both noise values, 100 steps
    from a uniform start against the oracle under the identical noise stream."""
    p = c2_params(D=D)
    x, y, th = oracle_mod.random_state(1, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    err = max(np.abs(gs["xu"] - os_["xu"]).max(), np.abs(gs["yu"] - os_["yu"]).max(),
              angle_diff(gs["theta"], os_["theta"]).max())
    assert err < TRAJ_TOL, err
    assert abs(e.polarization()[0] - o.polarization()[0]) < 1e-10
    assert e.info().tile_rebuilds > 0  # the headline path rebuilds through the tile sort


def test_trajectory_above_streaming_store_limit(oracle_mod):
    """300,000 particles (above the size up to which stage outputs use streaming stores) on a
    triangular lattice: 20 fused steps against the oracle."""
    n, L = 300000, 0.5 * math.sqrt(300000)
    p = c2_params(n=n, L=L, D=0.2)
    th = oracle_mod.random_state(3, n, L, L)[2]
    x, y = S.triangle_lattice(n, L, L)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.multi_step(20)
    o.multi_step(20)
    gs, os_ = e.get_state(), o.get_state()
    err = max(np.abs(gs["xu"] - os_["xu"]).max(), np.abs(gs["yu"] - os_["yu"]).max(),
              angle_diff(gs["theta"], os_["theta"]).max())
    assert err < TRAJ_TOL, err


@pytest.mark.parametrize("case", sorted(TRAJ_CASES))
def test_trajectory_100_steps(oracle_mod, case):
    p = TRAJ_CASES[case]()
    x, y, th = oracle_mod.random_state(1, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    err = max(np.abs(gs["xu"] - os_["xu"]).max(), np.abs(gs["yu"] - os_["yu"]).max(),
              angle_diff(gs["theta"], os_["theta"]).max())
    assert err < TRAJ_TOL, err
    gp, op = e.polarization()[0], o.polarization()[0]
    assert abs(gp - op) < 1e-10
    assert e.info().steps == 100


def test_reference_recipe_trajectory(oracle_mod):
    """RepulsiveParticle with the reference's own numbers (R12 truncated rc 1.05, R_a 1.1,
    dt 1/1024) and its fixed-period rebuild (cell_update_period = 32)."""
    p = ref_params(n=4096, Lx=32.0, Ly=32.0, rebuild=S.REBUILD_PERIOD, rebuild_period=16)
    x, y, th = _lattice(64, 32.0, 6)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL
    assert angle_diff(gs["theta"], os_["theta"]).max() < TRAJ_TOL


def test_quirk_mode_trajectory(oracle_mod):
    """Reference binning end to end: with a fixed rebuild period both sides use the same
    (quirk-pruned) lists, so the trajectories agree."""
    p = c2_params(n=8192, L=64.0, r_list=1.0625, ncell_x=113, ncell_y=113, binning=S.BIN_REF_QUIRK,
                  R_align=1.0, rc_rep=1.0, rebuild=S.REBUILD_PERIOD, rebuild_period=4, v0=0.5)
    x, y, th = oracle_mod.random_state(12, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL
    assert angle_diff(gs["theta"], os_["theta"]).max() < TRAJ_TOL


# ---------------------------------------------------------------------------------------------
# Engine mechanics
# ---------------------------------------------------------------------------------------------

def test_rebuild_schedules_agree(oracle_mod):
    """Displacement-triggered rebuilds with rollback give the same trajectory as rebuilding
    every step (any valid list gives the same forces)."""
    p = c2_params(n=16384, L=64.0)
    x, y, th = oracle_mod.random_state(8, p.n, p.Lx, p.Ly)
    a = S.Engine(p)
    a.set_state(x, y, th)
    a.multi_step(200)
    b = S.Engine(c2_params(n=16384, L=64.0, rebuild=S.REBUILD_PERIOD, rebuild_period=1))
    b.set_state(x, y, th)
    b.multi_step(200)
    sa, sb = a.get_state(), b.get_state()
    assert np.abs(sa["xu"] - sb["xu"]).max() < 1e-11
    assert angle_diff(sa["theta"], sb["theta"]).max() < 1e-11
    ia = a.info()
    assert 2 <= ia.rebuilds < 200
    assert ia.host_syncs < 200


def test_one_step_loop_rebuilds_only_when_due(oracle_mod):
    """A Box::One_Step loop (one spp_multi_step call per step) continues on the current list while
    the displacement allows it: 50 calls rebuild about as often as one 50-step call, far fewer than
    50 times, and follow the same trajectory (the oracle's within the trajectory bar)."""
    n, L = 16384, 64.0
    p = c2_params(n=n, L=L)
    th = oracle_mod.random_state(13, n, L, L)[2]
    x, y = S.triangle_lattice(n, L, L)  # no overlaps: displacements grow smoothly
    a = S.Engine(p)
    a.set_state(x, y, th)
    r0 = a.info().rebuilds
    a.multi_step(50)
    b = S.Engine(p)
    b.set_state(x, y, th)
    for _ in range(50):
        b.one_step()
    o = oracle_mod.Oracle(p)
    o.set_state(x, y, th)
    o.multi_step(50)
    ra, rb = a.info().rebuilds - r0, b.info().rebuilds - r0
    assert rb <= ra + 3 and rb < 12, (ra, rb)
    sa, sb, so = a.get_state(), b.get_state(), o.get_state()
    assert np.abs(sa["xu"] - sb["xu"]).max() < 1e-11
    assert angle_diff(sa["theta"], sb["theta"]).max() < 1e-11
    assert np.abs(sb["xu"] - so["xu"]).max() < TRAJ_TOL and angle_diff(sb["theta"], so["theta"]).max() < TRAJ_TOL
    assert b.info().steps == 50


def test_split_call_loop_with_caller_list_updates(oracle_mod):
    """Box::Interact + Move_Runge_Kutta2_1/2 for 60 steps, the caller refreshing the cells and the
    list every 10 steps (Update_Cells at the reference's cell_update_period cadence): Interact uses
    the list of the last build, as serial/box.h:322-371 does, so the trajectory stays on the
    oracle's driven the same way."""
    n, L = 8192, 48.0
    p = c2_params(n=n, L=L, noise_mode=S.NOISE_INJECTED)
    x, y, th = oracle_mod.random_state(19, n, L, L)
    e, o = _pair(p, x, y, th, oracle_mod)
    rng = np.random.default_rng(2)
    for step in range(60):
        if step and step % 10 == 0:
            e.build_neighbors()
            o.update_cells()
        dn = rng.normal(0, 0.02, n)
        e.inject_noise(dn)
        o.set_noise(dn)
        for st in (1, 2):
            e.interact()
            o.interact()
            e.move_stage(st)
            o.move_stage(st)
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL
    assert angle_diff(gs["theta"], os_["theta"]).max() < TRAJ_TOL


def test_snapshot_restore_replays_bitwise(oracle_mod):
    p = c2_params(n=16384, L=64.0)
    x, y, th = oracle_mod.random_state(10, p.n, p.Lx, p.Ly)
    e = S.Engine(p)
    e.set_state(x, y, th)
    e.multi_step(20)
    e.snapshot(3)
    e.multi_step(50)
    s1 = e.get_state()
    e.restore(3)
    e.multi_step(50)
    s2 = e.get_state()
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k


def test_list_capacity_grows(oracle_mod):
    """A clustered configuration overflows a small ELL capacity; the engine grows it and the
    list still equals the oracle's."""
    n, L = 4000, 16.0
    p = c2_params(n=n, L=L, ell_capacity=8)
    x, y, th = oracle_mod.random_state(14, n, L, L)
    x = x * 0.4  # squeeze into a 2.5x denser strip
    e, o = _pair(p, x, y, th, oracle_mod)
    assert e.info().ell_capacity > 8
    off, idx = e.get_neighbors()
    assert csr_pair_set(off, idx) == pair_set(*o.half_list())


def test_noise_stream_matches_oracle(oracle_mod):
    """Isolated particles (no neighbours): one RK2 step turns theta by exactly the noise draw."""
    n = 1000
    p = c2_params(n=n, L=200.0, D_rot=0.3, g=0.0, repulsion=S.REP_NONE, seed=99)
    x, y, th = oracle_mod.random_state(3, n, 200.0, 200.0)
    e, o = _pair(p, x, y, th, oracle_mod)
    e.one_step()
    o.one_step()
    amp = math.sqrt(2 * 0.3 * 0.005)
    want = np.array([amp * oracle_mod.noise_normal(99, i, 0) for i in range(n)])
    got = e.get_state()["theta"]
    assert angle_diff(got, th + want).max() < 1e-14
    assert angle_diff(o.get_state()["theta"], th + want).max() < 1e-14


def test_polarization_matches_oracle(oracle_mod):
    p = c2_params(n=20000, L=64.0)
    x, y, th = oracle_mod.random_state(15, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    gp, gx, gy = e.polarization()
    op, ox, oy = o.polarization()
    assert abs(gp - op) < 1e-13 and abs(gx - ox) < 1e-9 and abs(gy - oy) < 1e-9


# ---------------------------------------------------------------------------------------------
# Interaction-kernel variants: same physics, different gather strategies
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("variant", [S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP])
def test_kernel_variants_match_oracle(oracle_mod, variant):
    p = c2_params(n=16384, L=64.0, kernel_variant=variant)
    x, y, th = oracle_mod.random_state(21, p.n, p.Lx, p.Ly)
    e, o = _pair(p, x, y, th, oracle_mod)
    off, idx = e.get_neighbors()
    assert csr_pair_set(off, idx) == pair_set(*o.half_list())
    e.interact()
    o.interact()
    gf, of = e.get_forces(), o.get_forces()
    assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
    assert rel_err(gf[2], of[2]) < FORCE_TOL
    e.multi_step(100)
    o.multi_step(100)
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL
    assert angle_diff(gs["theta"], os_["theta"]).max() < TRAJ_TOL


@pytest.mark.parametrize("variant", [S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP])
def test_kernel_variants_dense_clusters(oracle_mod, variant):
    """Particles squeezed into 30 % of the box width: long lists and strong overlaps."""
    n, L = 6000, 16.0
    p = c2_params(n=n, L=L, kernel_variant=variant, dt=0.0005)  # strong overlaps: small steps
    x, y, th = oracle_mod.random_state(14, n, L, L)
    x = x * 0.3
    e, o = _pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gf, of = e.get_forces(), o.get_forces()
    assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
    assert rel_err(gf[2], of[2]) < FORCE_TOL
    e.multi_step(20)
    o.multi_step(20)
    gs, os_ = e.get_state(), o.get_state()
    assert np.abs(gs["xu"] - os_["xu"]).max() < TRAJ_TOL


@pytest.mark.parametrize("integ", [S.RK2, S.RK4])
def test_deep_variant_bitwise_equals_slot(oracle_mod, integ):
    """From 393,216 particles the slot kernel keeps four neighbour records in flight per lane and
    the deep variant eight; both add the pairs in list order, so the trajectories are bit-for-bit
    equal."""
    n, L = 409600, 320.0
    th = oracle_mod.random_state(8, n, L, L)[2]
    x, y = S.triangle_lattice(n, L, L)  # no deep overlaps on a lattice
    out = []
    for v in (S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP):
        e = S.Engine(c2_params(n=n, L=L, kernel_variant=v, D=0.2).update(integrator=integ))
        e.set_state(x, y, th)
        e.multi_step(60)
        out.append(e.get_state())
    for k in ("xw", "yw", "xu", "yu", "theta"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_set_state_rejects_positions_outside_the_box_and_non_finite_values(oracle_mod):
    """Wrapped positions must lie in [-L, L] and every value must be finite; a rejected state
    leaves the engine's previous state in place."""
    p = c1_params(n=256, Lx=8.0, Ly=8.0)
    x, y, th = oracle_mod.random_state(3, 256, 8.0, 8.0)
    e = S.Engine(p)
    e.set_state(x, y, th)
    before = e.get_state()
    for k, bad in (("x", 8.5), ("y", -9.0), ("x", np.nan), ("th", np.inf)):
        xb, yb, tb = x.copy(), y.copy(), th.copy()
        {"x": xb, "y": yb, "th": tb}[k][17] = bad
        with pytest.raises(S.SppError):
            e.set_state(xb, yb, tb)
    with pytest.raises(ValueError):  # the binding checks lengths before the C side reads n values
        e.set_state(x[:-1], y, th)
    edge = x.copy()
    edge[5] = 8.0  # the box edge itself is accepted
    e.set_state(edge, y, th)
    e.set_state(x, y, th)
    assert np.array_equal(e.get_state()["xw"], before["xw"])

"""GPU: single-bead active Brownian swimmers inside a ring membrane (bonds + bending).

Against the oracle: forces of one interaction within 1e-12, 100-step trajectories (RK2, Euler; the
vesicle at its full size of 1024 ring beads and 8192 swimmers) within 1e-9, the split-call path.
Closed forms: square-ring relaxation, chiral circle, and the free active-Brownian mean-square
displacement 2 v^2 / D^2 (D t - 1 + exp(-D t)) within its statistical error.
"""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff, vec_rel_err

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
TRAJ_TOL = 1e-9


def vesicle_params(nm, ns, R, **kw):
    """Soft (harmonic, k = 10 over radius 1) repulsion between all beads and swimmers; ring bonds of
    rest length 1 with k = 100 and bending kappa = 10; swimmers at speed 1 (F0 = 2, mobility 0.5)
    with rotational diffusion 0.1; RK2 dt = 1e-3."""
    p = S.default_params(S.MODEL_ABP)
    p.update(n=nm + ns, n_membrane=nm, Lx=R + 6, Ly=R + 6, integrator=S.RK2, dt=1e-3,
             repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, r_list=1.3, D_rot=0.1,
             v0=2.0, mob_par=0.5, mob_perp=0.5, mob_rot=1.0, k_bond=100.0, l_bond=1.0,
             kappa_bend=10.0, seed=6)
    return p.update(**kw)


def vesicle_state(nm, ns, R, seed):
    """Ring of nm beads at radius R, swimmers uniform in the disk of radius R - 1."""
    a = 2 * np.pi * np.arange(nm) / nm
    rng = np.random.default_rng(seed)
    r = (R - 1) * np.sqrt(rng.uniform(0, 1, ns))
    phi = rng.uniform(0, 2 * np.pi, ns)
    x = np.concatenate([R * np.cos(a), r * np.cos(phi)])
    y = np.concatenate([R * np.sin(a), r * np.sin(phi)])
    th = np.concatenate([np.zeros(nm), rng.uniform(-np.pi, np.pi, ns)])
    return x, y, th


def pair(p, x, y, th, oracle_mod):
    e = S.Engine(p)
    e.set_state(x, y, th)
    o = oracle_mod.Oracle(p)
    o.set_state(x, y, th)
    return e, o


def traj_err(e, o):
    g, w = e.get_state(), o.get_state()
    return max(np.abs(g["xu"] - w["xu"]).max(), np.abs(g["yu"] - w["yu"]).max(),
               angle_diff(g["theta"], w["theta"]).max())


@pytest.mark.parametrize("kappa", [0.0, 10.0])
def test_vesicle_forces_match_oracle(oracle_mod, kappa):
    nm, ns, R = 256, 2048, 0.5 / math.sin(math.pi / 256) * 1.05
    p = vesicle_params(nm, ns, R, kappa_bend=kappa)
    x, y, th = vesicle_state(nm, ns, R, 1)
    # jitter the ring so bending forces are not zero
    x[:nm] += 0.05 * np.sin(7 * np.arange(nm))
    e, o = pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gx, gy, gt = e.get_forces()
    ox, oy, ot = o.get_forces()
    assert vec_rel_err(gx, gy, ox, oy) < FORCE_TOL
    assert np.abs(gx[:nm] - ox[:nm]).max() < 1e-10 * max(1.0, np.abs(ox[:nm]).max())
    assert np.all(gt == p.torque0)


@pytest.mark.parametrize("integrator", [S.RK2, S.EULER])
def test_vesicle_trajectory_100_steps(oracle_mod, integrator):
    nm, ns = 512, 4096
    R = 0.5 / math.sin(math.pi / nm) * 1.1  # bonds pre-stretched by 10 %
    p = vesicle_params(nm, ns, R, integrator=integrator)
    x, y, th = vesicle_state(nm, ns, R, 2)
    e, o = pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    assert traj_err(e, o) < TRAJ_TOL


def test_vesicle_full_size_trajectory(oracle_mod):
    """1024 ring beads at radius 180 (bonds pre-stretched to 1.10) and 8192 swimmers inside."""
    nm, ns, R = 1024, 8192, 180.0
    p = vesicle_params(nm, ns, R)
    x, y, th = vesicle_state(nm, ns, R, 3)
    e, o = pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    assert traj_err(e, o) < TRAJ_TOL
    assert e.info().steps == 100


def test_split_call_path_matches_oracle(oracle_mod):
    nm, ns = 128, 1024
    R = 0.5 / math.sin(math.pi / nm) * 1.05
    p = vesicle_params(nm, ns, R, noise_mode=S.NOISE_INJECTED)
    x, y, th = vesicle_state(nm, ns, R, 4)
    e, o = pair(p, x, y, th, oracle_mod)
    rng = np.random.default_rng(9)
    for _ in range(5):
        d = 0.02 * rng.standard_normal(p.n)
        e.inject_noise(d)
        o.set_noise(d)
        for st in (1, 2):
            e.interact()
            e.move_stage(st)
            o.interact()
            o.move_stage(st)
    assert traj_err(e, o) < TRAJ_TOL


def test_square_ring_relaxation_closed_form():
    n, s0, l0, K, m, dt, steps = 4, 1.3, 1.0, 100.0, 0.5, 1e-4, 200
    R0 = s0 / (2 * math.sin(math.pi / n))
    p = S.default_params(S.MODEL_ABP).update(
        n=n, n_membrane=n, Lx=4.0, Ly=4.0, k_bond=K, l_bond=l0, dt=dt, kappa_bend=10.0, mob_par=m,
        mob_perp=m, repulsion=S.REP_NONE, r_list=1.3, D_rot=0.0, integrator=S.RK2)
    a = 0.3 + 2 * np.pi * np.arange(n) / n
    e = S.Engine(p)
    e.set_state(R0 * np.cos(a), R0 * np.sin(a), np.zeros(n))
    e.multi_step(steps)
    st = e.get_state()
    side = np.hypot(np.diff(np.append(st["xu"], st["xu"][0])), np.diff(np.append(st["yu"], st["yu"][0])))
    gamma = 4 * m * K * math.sin(math.pi / n) ** 2
    assert np.allclose(side, l0 + (s0 - l0) * math.exp(-gamma * steps * dt), rtol=1e-5)


def test_chiral_swimmers_closed_form():
    h, F0, m, k, tq, n = 0.01, 2.0, 0.5, 1.0, 0.8, 64
    p = S.default_params(S.MODEL_ABP).update(
        n=n, n_membrane=0, Lx=40.0, Ly=40.0, v0=F0, mob_par=m, mob_perp=m, mob_rot=k, torque0=tq,
        dt=h, repulsion=S.REP_NONE, r_list=1.3, D_rot=0.0, integrator=S.RK2)
    rng = np.random.default_rng(3)
    x0, y0, th0 = rng.uniform(-30, 30, n), rng.uniform(-30, 30, n), rng.uniform(-np.pi, np.pi, n)
    e = S.Engine(p)
    e.set_state(x0, y0, th0)
    w, v = k * tq, m * F0
    Rd = h * v / (2 * math.sin(w * h / 2))
    cx, cy = x0 - Rd * np.sin(th0), y0 + Rd * np.cos(th0)
    for _ in range(10):
        e.multi_step(17)
        st = e.get_state()
        assert np.abs(np.hypot(st["xu"] - cx, st["yu"] - cy) - Rd).max() < 1e-10


def test_free_abp_mean_square_displacement():
    """Non-interacting swimmers: MSD(t) = 2 v^2/D^2 (D t - 1 + exp(-D t)), v = mob_par F0."""
    n, D, F0, m, dt = 40000, 0.5, 2.0, 0.5, 0.01
    L = 0.5 * math.sqrt(n / 0.5)
    p = S.default_params(S.MODEL_ABP).update(
        n=n, n_membrane=0, Lx=L, Ly=L, v0=F0, mob_par=m, mob_perp=m, D_rot=D, dt=dt,
        repulsion=S.REP_NONE, r_list=1.3, integrator=S.RK2, seed=17)
    x, y, th = S.random_state(4, n, L, L)
    e = S.Engine(p)
    e.set_state(x, y, th)
    v, t = m * F0, 0.0
    for steps in (100, 200, 400):
        e.multi_step(steps)
        t += steps * dt
        st = e.get_state()
        sq = (st["xu"] - x) ** 2 + (st["yu"] - y) ** 2
        want = 2 * v * v / D ** 2 * (D * t - 1 + math.exp(-D * t))
        err = sq.std() / math.sqrt(n)
        assert abs(sq.mean() - want) < 4 * err + 2e-3 * want, (t, sq.mean(), want, err)


def test_abp_rejects_unsupported():
    p = vesicle_params(16, 16, 10.0)
    with pytest.raises(S.SppError):
        S.Engine(p.copy().update(integrator=S.RK4))
    with pytest.raises(S.SppError):
        S.Engine(p.copy().update(n_membrane=2))
    with pytest.raises(S.SppError):
        S.Engine(p.copy().update(kernel_variant=S.KERNEL_SLOT_DEEP))  # the slot kernel only
    with pytest.raises(S.SppError):
        S.Engine(p, [S.system(32, 16.0)])


# ---------------------------------------------------------------------------------------------
# Chains of several beads (nb > 1) inside a ring membrane
# ---------------------------------------------------------------------------------------------

def chain_vesicle_params(nm, ns, R, nb, **kw):
    """Ring beads with harmonic mutual repulsion and bonds (k = 100, bending 10); swimmers are chains
    of nb beads one unit apart with the reference's bead law (R12 truncated, sigma 1, cutoff 1.1,
    amplitude 1) and Set_nb mobilities; RK2 dt = 1e-3, rotational diffusion 0.1."""
    p = S.default_params(S.MODEL_ABP)
    p.update(n=nm + ns, n_membrane=nm, Lx=R + 6, Ly=R + 6, integrator=S.RK2, dt=1e-3,
             repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, D_rot=0.1, v0=2.0,
             mob_par=0.5, mob_perp=0.5, mob_rot=1.0, k_bond=100.0, l_bond=1.0, kappa_bend=10.0, seed=11,
             nb_swimmer=nb, bead_sep=1.0, sigma_bead=1.0, rc_bead=1.1, A_bead=1.0,
             r_list=1.1 + (nb - 1) * 1.0 + 0.4)
    return p.update(**kw)


def chain_lattice_state(nm, R, spacing, seed):
    """Ring of nm beads at radius R (bonds at rest length 1 for R = 0.5 / sin(pi / nm)); chains on a
    square lattice inside radius R - spacing / 2, random headings (no initial bead overlaps)."""
    a = 2 * np.pi * np.arange(nm) / nm
    g = np.arange(-R, R + 1e-9, spacing)
    gx, gy = [v.ravel() for v in np.meshgrid(g, g)]
    keep = np.hypot(gx, gy) < R - spacing / 2
    gx, gy = gx[keep], gy[keep]
    rng = np.random.default_rng(seed)
    x = np.concatenate([R * np.cos(a), gx + rng.uniform(-0.2, 0.2, gx.size)])
    y = np.concatenate([R * np.sin(a), gy + rng.uniform(-0.2, 0.2, gy.size)])
    th = np.concatenate([np.zeros(nm), rng.uniform(-np.pi, np.pi, gx.size)])
    return x, y, th


@pytest.mark.parametrize("nb", [2, 3])
def test_chain_forces_match_oracle(oracle_mod, nb):
    """One interaction of a ring plus randomly placed, overlapping chains: forces and torques
    (bead-pair torques s x F included) within 1e-12 of the oracle."""
    nm, ns = 128, 700
    R = 0.5 / math.sin(math.pi / nm)
    p = chain_vesicle_params(nm, ns, R, nb)
    x, y, th = vesicle_state(nm, ns, R, 5)
    e, o = pair(p, x, y, th, oracle_mod)
    e.interact()
    o.interact()
    gf, of = e.get_forces(), o.get_forces()
    assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
    tau_err = np.linalg.norm(gf[2] - of[2]) / np.linalg.norm(of[2])
    assert np.linalg.norm(of[2][nm:]) > 0 and tau_err < FORCE_TOL, tau_err


@pytest.mark.parametrize("nb,integ", [(2, S.RK2), (3, S.RK2), (3, S.EULER)])
def test_chain_trajectory_matches_oracle(oracle_mod, nb, integ):
    """100 steps of chains swimming inside a bending ring, against the oracle under the identical
    noise stream, within 1e-9."""
    nm = 256
    R = 0.5 / math.sin(math.pi / nm)
    x, y, th = chain_lattice_state(nm, R, nb + 0.6, 9)
    p = chain_vesicle_params(nm, x.size - nm, R, nb, integrator=integ, v0=5.0)
    e, o = pair(p, x, y, th, oracle_mod)
    e.multi_step(100)
    o.multi_step(100)
    assert traj_err(e, o) < TRAJ_TOL


def test_chain_split_stages_match_oracle(oracle_mod):
    """Box::Interact + Move_Runge_Kutta2_1/2 one call at a time with chains: every stage's forces,
    torques and state against the oracle."""
    nm = 128
    R = 0.5 / math.sin(math.pi / nm)
    x, y, th = chain_lattice_state(nm, R, 2.6, 4)
    p = chain_vesicle_params(nm, x.size - nm, R, 2, v0=5.0, noise_mode=S.NOISE_INJECTED)
    e, o = pair(p, x, y, th, oracle_mod)
    rng = np.random.default_rng(3)
    for _ in range(3):
        dn = rng.normal(0, 0.05, p.n)
        e.inject_noise(dn)
        o.set_noise(dn)
        for st in (1, 2):
            e.interact()
            o.interact()
            gf, of = e.get_forces(), o.get_forces()
            assert vec_rel_err(gf[0], gf[1], of[0], of[1]) < FORCE_TOL
            assert np.abs(gf[2] - of[2]).max() <= FORCE_TOL * max(1.0, np.abs(of[2]).max())
            e.move_stage(st)
            o.move_stage(st)
    assert traj_err(e, o) < 1e-12

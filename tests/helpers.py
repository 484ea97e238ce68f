"""Shared configurations and comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

import spp_amd as S


def c1_params(**kw):
    """Continuous-time Vicsek: N=4096, side 64 (Lx=32), R=1, v0=0.5, g=1, D=0.1, RK2, dt=0.01."""
    p = S.default_params(S.MODEL_CONTINUOUS)
    p.update(n=4096, Lx=32.0, Ly=32.0, dt=0.01, v0=0.5, g=1.0, R_align=1.0, alpha=0.0,
             D_rot=0.1, integrator=S.RK2, r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT, seed=1)
    return p.update(**kw)


def c2_params(n=262144, L=256.0, D=0.05, **kw):
    """Repulsive Vicsek: density 1, v0=1, R_a=1, harmonic k=10 over rc=1, RK2, dt=0.005,
    list radius 1.3 (skin 0.3), rebuilt on displacement."""
    p = S.default_params(S.MODEL_REPULSIVE)
    p.update(n=n, Lx=L, Ly=L, dt=0.005, v0=1.0, g=1.0, R_align=1.0, repulsion=S.REP_HARMONIC,
             A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, D_rot=D, integrator=S.RK2, r_list=1.3,
             rebuild=S.REBUILD_DISPLACEMENT, seed=2)
    return p.update(**kw)


def ref_params(**kw):
    """The reference's own RepulsiveParticle recipe (parallel/repulsive-vicsek.cpp:133-139):
    sigma 1, rc 1.05, R_a 1.1, R12 truncated, reference cell grid and rv."""
    p = S.default_params(S.MODEL_REPULSIVE)
    p.update(n=8192, Lx=64.0, Ly=64.0, repulsion=S.REP_R12_TRUNCATED, A_rep=2.0, sigma_rep=1.0,
             rc_rep=1.05, R_align=1.1, g=1.0, dt=1.0 / 1024, r_list=1.1625,
             D_rot=0.1, integrator=S.RK2, seed=3)
    return p.update(**kw)


def rel_err(a, b):
    """Norm-wise relative error ||a - b|| / ||b|| (per-particle relative error is ill-posed when
    pair sums nearly cancel)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def vec_rel_err(ax, ay, bx, by):
    """||(ax, ay) - (bx, by)|| / ||(bx, by)|| over all particles (2-vectors per particle)."""
    num = np.sqrt(np.sum((np.asarray(ax) - bx) ** 2 + (np.asarray(ay) - by) ** 2))
    den = np.sqrt(np.sum(np.asarray(bx) ** 2 + np.asarray(by) ** 2))
    return num / (den if den > 0 else 1.0)


def angle_diff(a, b):
    d = np.asarray(a) - np.asarray(b)
    return np.abs(d - 2 * np.pi * np.round(d / (2 * np.pi)))


def pair_set(owner, other):
    lo = np.minimum(owner, other).astype(np.int64)
    hi = np.maximum(owner, other).astype(np.int64)
    return set((lo << 32 | hi).tolist())


def csr_pair_set(off, idx):
    n = off.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    return pair_set(rows, idx)

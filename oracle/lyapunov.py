"""CPU oracle of the finite-difference Lyapunov step (numpy restatement of the reference).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the engine's tangent-vector code.

Reference algorithm (parallel/lyapunovbox.h:128-194, Evolution_Reorthonormalize):
  * every direction has a perturbed clone of the state that sees the same noise;
  * after an interval the deviation is clone - base, with State_Hyper_Vector's differences:
    positions through C2DVector::Periodic_Transform (shared/c2dvector.h:67-73) and theta wrapped
    by theta - 2 pi ceil((theta - pi)/(2 pi)) (shared/state-hyper-vector.h:96-106, 213-225);
  * the inner product is sum(r.r' + theta theta') (:157-167);
  * the deviations are Gram-Schmidt'ed in index order, u_i -= u_j (u_j . u_i) for j < i, then
    u_i /= |u_i| (shared/vector-set.h:118-127);
  * the growth of direction i is (v_i . u_i) / amplitude (lyapunovbox.h:177-183);
  * the clones restart at base + scaled u_i (Add_Deviation, lyapunovbox.h:47-67: r += dr then
    Periodic_Transform, theta += dtheta, v = (cos, sin)).
The reference keeps the projected deviation unrescaled (cumulative ratio); `rescale=True` restarts
every clone at the initial amplitude instead (Benettin).  Either way the accumulated
log(growth / start length) is the same in the linear regime.
"""
from __future__ import annotations

import numpy as np


def ref_wrap(x, L2):
    """C2DVector::Periodic_Transform component: x - L2*(int)floor(x/L2 + 0.5) (c2dvector.h:67-73)."""
    x = np.asarray(x, dtype=np.float64)
    return x - L2 * np.trunc(np.floor(x / L2 + 0.5))


def wrap_theta_diff(t):
    """theta - 2 pi ceil((theta - pi)/(2 pi)), into (-pi, pi] (state-hyper-vector.h:223)."""
    t = np.asarray(t, dtype=np.float64)
    return t - 2 * np.pi * np.ceil((t - np.pi) / (2 * np.pi))


def deviation(base: dict, clone: dict, Lx: float, Ly: float) -> np.ndarray:
    """clone - base as a [3, n] array (x, y, theta) (State_Hyper_Vector::operator-)."""
    return np.stack([
        ref_wrap(clone["xw"] - base["xw"], 2 * Lx),
        ref_wrap(clone["yw"] - base["yw"], 2 * Ly),
        wrap_theta_diff(clone["theta"] - base["theta"]),
    ])


def gram_schmidt(V: np.ndarray):
    """VectorSet::Renormalize(us) over all K rows of V ([K, ...] arrays): returns (U, growth) with
    U orthonormal in index order and growth_i = V_i . U_i (= the QR diagonal R_ii)."""
    K = V.shape[0]
    flat = V.reshape(K, -1)
    U = np.empty_like(flat)
    for i in range(K):
        u = flat[i].copy()
        for j in range(i):
            u -= U[j] * np.dot(U[j], u)
        U[i] = u / np.sqrt(np.dot(u, u))
    growth = np.array([np.dot(flat[i], U[i]) for i in range(K)])
    return U.reshape(V.shape), growth


def add_deviation(base: dict, d: np.ndarray, Lx: float, Ly: float) -> dict:
    """LyapunovBox::Add_Deviation (lyapunovbox.h:47-67) on a state dict; the unwrapped position moves
    with the wrapped one (the engine's RK stages rebuild r from it)."""
    return {
        "xw": ref_wrap(base["xw"] + d[0], 2 * Lx),
        "yw": ref_wrap(base["yw"] + d[1], 2 * Ly),
        "xu": base["xu"] + d[0],
        "yu": base["yu"] + d[1],
        "theta": base["theta"] + d[2],
    }


class LyapunovOracle:
    """One replica: a base oracle system plus K clones, each an oracle.Oracle of the same params."""

    def __init__(self, systems, Lx: float, Ly: float, eps: float, rescale: bool = True):
        self.base, self.clones = systems[0], list(systems[1:])
        self.K = len(self.clones)
        self.Lx, self.Ly, self.eps, self.rescale = Lx, Ly, eps, rescale
        self.norm0 = np.full(self.K, eps)
        self.logsum = np.zeros(self.K)
        self.U = None

    def multi_step(self, m: int):
        self.base.multi_step(m)
        for c in self.clones:
            c.multi_step(m)

    def renormalize(self):
        b = self.base.get_state()
        V = np.stack([deviation(b, c.get_state(), self.Lx, self.Ly) for c in self.clones])
        U, growth = gram_schmidt(V)
        self.logsum += np.log(growth / self.norm0)
        scale = np.full(self.K, self.eps) if self.rescale else growth
        self.norm0 = scale.copy()
        for k, c in enumerate(self.clones):
            s = add_deviation(b, scale[k] * U[k], self.Lx, self.Ly)
            c.set_state(s["xw"], s["yw"], s["theta"], s["xu"], s["yu"])
        self.U = U
        return self.logsum.copy()

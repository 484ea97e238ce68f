"""Parameter sweeps over the order-disorder transition of the continuous-time Vicsek model.

Three ways to produce a polarisation-versus-noise curve P(D) at fixed density, speed and size,
every point resolved to a given error bar (the reference's autocorrelation estimator,
polarization-analysis.py:29-62, restated in observables.polarity_error):

* ``plain_curve``   -- the reference's way: one simulation per point on a uniform grid of D, run
                       until its error bar is small enough (serial/mpi-main.cpp:157-160 runs such
                       points as independent processes);
* ``batched_curve`` -- the same points as systems of one batched engine (one kernel launch advances
                       every point);
* ``steered_curve`` -- the batched engine on points placed by the surrogate network: the network,
                       trained on simulator output (``make_dataset``), predicts the curve; a grid
                       is refined where the predicted values jump, so the refinement a uniform
                       grid needs extra sequential rounds for is simulated at once, and points of
                       high predicted variance run as several replicas side by side.  Every
                       reported value is a simulated one; where neighbouring simulated values
                       still differ by more than the resolution, the sweep adds simulated
                       midpoints (the plain fallback).

The curve counts as resolved when every point's error bar is at most ``tol`` and neighbouring
points differ by at most ``gap`` (default 3 tol: two values each within tol of the truth can differ
by 2 tol through noise alone) or are already ``dD_min`` apart (the finest noise resolution asked
for: in a finite system the transition can be a jump, which no spacing resolves).  Runs start from
aligned headings (see ``initial_state``).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import (MODEL_CONTINUOUS, REBUILD_DISPLACEMENT, RK2, Engine, default_params, random_state, system)
from .observables import polarity_error

FEATURES = ("D", "rho", "v0", "logN")
TARGETS = ("mean_P", "var_P", "lambda1")


def model_params(dt=0.01):
    """Continuous-time Vicsek: unit interaction radius, g = 1, RK2, Verlet radius 1.3."""
    p = default_params(MODEL_CONTINUOUS)
    return p.update(dt=dt, R_align=1.0, g=1.0, alpha=0.0, integrator=RK2, r_list=1.3,
                    rebuild=REBUILD_DISPLACEMENT, device_timing=1)


def half_width(n, rho):
    return 0.5 * math.sqrt(n / rho)


def features(D, rho, v0, n):
    """Network inputs, scaled to order one."""
    return np.stack([4.0 * np.asarray(D) - 1.0, np.log(np.asarray(rho)), 2.0 * np.asarray(v0) - 1.0,
                     np.log(np.asarray(n, dtype=np.float64)) - 7.0], axis=-1).astype(np.float32)


@dataclass
class Timer:
    """Cost of engine calls: wall-clock seconds, device seconds (the engine's own HIP-event
    accounting of its stepping calls, spp_info.device_ms) and particle-steps."""
    seconds: float = 0.0
    device_seconds: float = 0.0
    particle_steps: float = 0.0

    def run(self, fn, *a, particles=0, steps=0):
        eng = getattr(fn, "__self__", None)
        d0 = eng.info().device_ms if isinstance(eng, Engine) else None
        t0 = time.perf_counter()
        r = fn(*a)
        self.seconds += time.perf_counter() - t0
        if d0 is not None:
            self.device_seconds += (eng.info().device_ms - d0) * 1e-3
        self.particle_steps += particles * steps
        return r


# -------------------------------------------------------------------------------------------------
# Simulation of many points in one batched engine
# -------------------------------------------------------------------------------------------------

def initial_state(seed, n, L, ordered):
    """Uniform positions; headings uniform, or all equal (`ordered`: the ordered phase is then
    reached at once and the disordered one within a few rotational-diffusion times)."""
    x, y, t = random_state(seed, int(n), L, L)
    if ordered:
        t = np.zeros_like(t)
    return x, y, t


def run_points(points, steps, sample_every, seed, timer=None, dt=0.01, state=None, ordered=True):
    """Simulate points [(D, rho, v0, n)] as one batched engine for `steps` steps, sampling every
    system's polarisation every `sample_every` steps.  Returns (series [samples, points], final
    state dict per point) ; `state` continues from a previous call's final states."""
    timer = timer or Timer()
    p = model_params(dt)
    systems = [system(int(n), half_width(n, rho), v0=v0, g=1.0, D_rot=D, seed=seed + 7919 * k)
               for k, (D, rho, v0, n) in enumerate(points)]
    e = Engine(p, systems)
    xs, ys, ts = [], [], []
    for k, (D, rho, v0, n) in enumerate(points):
        if state is not None:
            s = state[k]
            xs.append(s["xw"])
            ys.append(s["yw"])
            ts.append(s["theta"])
        else:
            L = half_width(n, rho)
            x, y, t = initial_state(seed * 1000003 + k, n, L, ordered)
            xs.append(x)
            ys.append(y)
            ts.append(t)
    timer.run(e.set_state, np.concatenate(xs), np.concatenate(ys), np.concatenate(ts))
    ntot = int(sum(n for *_, n in points))
    steps -= steps % sample_every
    series = timer.run(e.multi_step_sampled, steps, sample_every, particles=ntot, steps=steps)
    st = e.get_state()
    off = e.offsets
    final = [{k: v[off[i]:off[i + 1]] for k, v in st.items()} for i in range(len(points))]
    e.close()
    return series, final


def run_lyapunov(points, warm, steps, seed, sample_every=10, interval=10, eps=1e-6, dt=0.01, timer=None,
                 ordered=True):
    """Polarisation series and largest Lyapunov exponent of each point (one base + one clone per
    point, all points of one size n in one engine): `warm` steps of the pairs, then `steps` steps
    with renormalisation every `interval` steps and a polarisation sample of every base every
    `sample_every` steps (a multiple of `interval`).  Returns (series [samples, points], lambda)."""
    timer = timer or Timer()
    ns = {int(n) for *_, n in points}
    if len(ns) != 1:
        raise ValueError("one system size per Lyapunov batch")
    if sample_every % interval:
        raise ValueError("sample_every must be a multiple of interval")
    n = ns.pop()
    p = model_params(dt)
    systems = []
    xs, ys, ts = [], [], []
    for k, (D, rho, v0, _) in enumerate(points):
        L = half_width(n, rho)
        x, y, t = initial_state(seed * 1000003 + k, n, L, ordered)
        for _ in range(2):
            systems.append(system(n, L, v0=v0, g=1.0, D_rot=D, seed=seed + 7919 * k))
            xs.append(x)
            ys.append(y)
            ts.append(t)
    e = Engine(p, systems)
    timer.run(e.set_state, np.concatenate(xs), np.concatenate(ys), np.concatenate(ts))
    ntot = 2 * n * len(points)
    timer.run(e.multi_step, warm, particles=ntot, steps=warm)
    e.lyapunov_init(len(points), 1, eps, True, seed)
    series = []
    for _ in range(steps // sample_every):
        timer.run(e.lyapunov_run, sample_every, interval, particles=ntot, steps=sample_every)
        series.append(timer.run(e.batch_polarization)[0::2])
    sums, T = e.lyapunov_get()
    e.close()
    return np.array(series), sums[:, 0] / T


# -------------------------------------------------------------------------------------------------
# Training data and the surrogate
# -------------------------------------------------------------------------------------------------

@dataclass
class Dataset:
    x: np.ndarray                    # [n, 4] raw parameters (D, rho, v0, n)
    y: np.ndarray                    # [n, 3] mean P, var P, lambda1
    seconds: float = 0.0

    def inputs(self):
        return features(self.x[:, 0], self.x[:, 1], self.x[:, 2], self.x[:, 3])


def make_dataset(count, sizes=(256, 512, 1024), warm=6000, steps=6000, sample_every=20, seed=1,
                 D_range=(0.0, 1.5), rho_range=(0.25, 4.0), v0_range=(0.25, 1.0), per_batch=64):
    """Seeded parameter draws simulated in batches of `per_batch` points of one size each: mean and
    variance of the polarisation over the `steps` after `warm` and the largest exponent."""
    rng = np.random.default_rng(seed)
    timer = Timer()
    xs, ys = [], []
    k = 0
    while k < count:
        n = int(sizes[(k // per_batch) % len(sizes)])
        m = min(per_batch, count - k)
        pts = [(rng.uniform(*D_range), rng.uniform(*rho_range), rng.uniform(*v0_range), n) for _ in range(m)]
        series, lam = run_lyapunov(pts, warm, steps, seed + 31 * k, sample_every=sample_every, timer=timer)
        for j, pt in enumerate(pts):
            xs.append(pt)
            ys.append((series[:, j].mean(), series[:, j].var(), lam[j]))
        k += m
    return Dataset(np.array(xs, dtype=np.float64), np.array(ys, dtype=np.float64), timer.seconds)


@dataclass
class Surrogate:
    """The network plus the target scaling it was trained with."""
    net: object
    mu: np.ndarray
    sd: np.ndarray
    train_ms: float = 0.0

    def predict(self, D, rho, v0, n):
        y = self.net.forward(features(D, rho, v0, n)).astype(np.float64)
        return y * self.sd + self.mu


def train_surrogate(data: Dataset, epochs=400, batch=256, seed=1):
    from .surrogate import Mlp

    y = data.y.copy()
    y[:, 1] = np.log(np.maximum(y[:, 1], 1e-8))  # variance spans decades: fit its logarithm
    mu, sd = y.mean(axis=0), y.std(axis=0) + 1e-12
    net = Mlp(seed=seed, max_batch=max(batch, 4096))
    net.train(data.inputs(), ((y - mu) / sd).astype(np.float32), epochs=epochs, batch=batch, seed=seed)
    s = Surrogate(net, mu, sd, net.last_train_ms())
    return s


def merge(a: Dataset, b: Dataset) -> Dataset:
    return Dataset(np.concatenate([a.x, b.x]), np.concatenate([a.y, b.y]), a.seconds + b.seconds)


def simulate_points(points, warm=6000, steps=6000, sample_every=20, seed=1, per_batch=64):
    """Plain simulations of arbitrary (D, rho, v0, n) points, grouped by size into Lyapunov-pair
    batches as make_dataset runs them: mean and variance of P and the largest exponent."""
    timer = Timer()
    xs, ys = [], []
    by_n = {}
    for pt in points:
        by_n.setdefault(int(pt[3]), []).append(tuple(pt))
    k = 0
    for n, pts in sorted(by_n.items()):
        for b in range(0, len(pts), per_batch):
            chunk = pts[b:b + per_batch]
            series, lam = run_lyapunov(chunk, warm, steps, seed + 31 * k, sample_every=sample_every, timer=timer)
            for j, pt in enumerate(chunk):
                xs.append(pt)
                ys.append((series[:, j].mean(), series[:, j].var(), lam[j]))
            k += len(chunk)
    return Dataset(np.array(xs, dtype=np.float64).reshape(-1, 4), np.array(ys, dtype=np.float64).reshape(-1, 3),
                   timer.seconds)


def active_round(s: Surrogate, data: Dataset, per_round=128, candidates=4096, seed=1, sizes=(256, 512, 1024),
                 D_range=(0.0, 1.5), rho_range=(0.25, 4.0), v0_range=(0.25, 1.0), band=(0.15, 0.85), warm=6000,
                 steps=6000, sample_every=20, epochs=400, batch=256):
    """One round of variance-steered sampling: of `candidates` seeded draws of (D, rho, v0, N),
    the `per_round` with the highest predicted polarisation variance among those whose predicted
    mean polarisation lies in `band` (the transition; topped up from outside it by variance when
    the band holds too few) are simulated plainly and added to the training set, and the network
    is retrained.  Returns (surrogate, enlarged data set, the picked points, their predicted
    variances, the candidates' predicted variances)."""
    rng = np.random.default_rng(seed)
    cand = np.stack([rng.uniform(*D_range, candidates), rng.uniform(*rho_range, candidates),
                     rng.uniform(*v0_range, candidates), rng.choice(np.asarray(sizes, dtype=np.float64), candidates)],
                    axis=1)
    y = s.predict(cand[:, 0], cand[:, 1], cand[:, 2], cand[:, 3])
    P, var = y[:, 0], np.exp(y[:, 1])
    in_band = (P >= band[0]) & (P <= band[1])
    order = np.lexsort((-var, ~in_band))  # in-band first, each group by decreasing variance
    pick = order[:per_round]
    new = simulate_points([tuple(r) for r in cand[pick]], warm=warm, steps=steps, sample_every=sample_every,
                          seed=seed + 17)
    data2 = merge(data, new)
    s2 = train_surrogate(data2, epochs=epochs, batch=batch, seed=seed)
    return s2, data2, cand[pick], var[pick], var


def held_out_error(s: Surrogate, held: Dataset):
    """Root-mean-square error of the predicted mean polarisation against plain simulations."""
    pred = s.predict(held.x[:, 0], held.x[:, 1], held.x[:, 2], held.x[:, 3])[:, 0]
    return float(np.sqrt(np.mean((pred - held.y[:, 0]) ** 2)))


def predicted_curve(s: Surrogate, D, rho, v0, n):
    y = s.predict(D, np.full_like(D, rho), np.full_like(D, v0), np.full_like(D, n))
    return y[:, 0], np.exp(y[:, 1]), y[:, 2]


# -------------------------------------------------------------------------------------------------
# Curves
# -------------------------------------------------------------------------------------------------

@dataclass
class Curve:
    D: np.ndarray
    P: np.ndarray
    err: np.ndarray
    seconds: float                   # wall clock of the engine calls
    particle_steps: float
    points_simulated: int
    rounds: int = 1
    log: list = field(default_factory=list)
    device_seconds: float = 0.0      # device time of the engine calls

    def resolved(self, tol, gap, dD_min):
        o = np.argsort(self.D)
        ok_gap = (np.abs(np.diff(self.P[o])) <= gap) | (np.diff(self.D[o]) <= 2 * dD_min)
        return bool(np.all(self.err <= tol) and np.all(ok_gap))


def _converge(points, tol, chunk, sample_every, max_steps, seed, timer, min_samples=40, reps=None):
    """Run `points` in one batch until every error bar is <= tol.  Point i runs as reps[i]
    independent replicas (other noise keys and initial positions); its value is the mean of the
    replicas' means and its error bar sqrt(sum of squared replica errors) / replicas.  A point whose
    error bar is small enough leaves the batch at the next chunk boundary.  Returns (mean, error)."""
    n = len(points)
    reps = [1] * n if reps is None else [max(1, int(r)) for r in reps]
    owner = [i for i in range(n) for _ in range(reps[i])]
    series = [np.zeros(0) for _ in owner]
    state = None
    active = list(range(len(owner)))  # systems
    done_mean = np.zeros(n)
    done_err = np.full(n, np.inf)
    steps = 0
    seg = 0

    def estimate(i):
        ms, es = [], []
        for q, o in enumerate(owner):
            if o == i:
                m, e, _ = polarity_error(np.arange(series[q].size), series[q])
                ms.append(m)
                es.append(e)
        return float(np.mean(ms)), float(np.sqrt(np.sum(np.square(es))) / len(es))

    while active and steps < max_steps:
        pts = [points[owner[q]] for q in active]
        # a new segment re-keys the noise so it does not repeat the previous segment's stream
        s, state = run_points(pts, chunk, sample_every, seed + 104729 * seg, timer, state=state)
        seg += 1
        steps += chunk
        for j, q in enumerate(active):
            series[q] = np.concatenate([series[q], s[:, j]])
        finished = set()
        if series[active[0]].size >= min_samples:
            for i in {owner[q] for q in active}:
                done_mean[i], done_err[i] = estimate(i)
                if done_err[i] <= tol:
                    finished.add(i)
        keep = [(q, st) for q, st in zip(active, state) if owner[q] not in finished]
        active = [q for q, _ in keep]
        state = [st for _, st in keep]
    for i in {owner[q] for q in active}:  # out of steps: report what was measured
        done_mean[i], done_err[i] = estimate(i)
    return done_mean, done_err


def plain_curve(D_grid, rho, v0, n, tol=0.02, chunk=2000, sample_every=10, max_steps=200000, seed=1):
    """One engine per point, sequentially (the reference's independent runs)."""
    timer = Timer()
    P, E = [], []
    for k, D in enumerate(D_grid):
        m, e = _converge([(D, rho, v0, n)], tol, chunk, sample_every, max_steps, seed + 11 * k, timer)
        P.append(m[0])
        E.append(e[0])
    return Curve(np.array(D_grid), np.array(P), np.array(E), timer.seconds, timer.particle_steps, len(D_grid),
                 device_seconds=timer.device_seconds)


def _refine(D, P, E, points_fn, tol, gap, dD_min, chunk, sample_every, max_steps, seed, timer, rounds, log,
            max_rounds, reps_fn=None):
    """Add simulated midpoints between neighbours that differ by more than `gap` and are more than
    2 dD_min apart (`reps_fn(D)`: replicas per new point)."""
    while rounds < max_rounds:
        o = np.argsort(D)
        D, P, E = D[o], P[o], E[o]
        bad = np.nonzero((np.abs(np.diff(P)) > gap) & (np.diff(D) > 2 * dD_min))[0]
        if bad.size == 0:
            break
        mids = 0.5 * (D[bad] + D[bad + 1])
        reps = reps_fn(mids) if reps_fn else None
        m, e = _converge(points_fn(mids), tol, chunk, sample_every, max_steps, seed + 1009 * rounds, timer, reps=reps)
        log.append(("refine", len(mids)))
        D, P, E = np.concatenate([D, mids]), np.concatenate([P, m]), np.concatenate([E, e])
        rounds += 1
    o = np.argsort(D)
    return D[o], P[o], E[o], rounds


def batched_curve(D_grid, rho, v0, n, tol=0.02, gap=None, dD_min=None, chunk=2000, sample_every=10,
                  max_steps=200000, seed=1, max_rounds=6):
    """All points of a uniform grid in one batched engine, refined where neighbours differ by more
    than `gap` (no network)."""
    gap = 3 * tol if gap is None else gap
    D = np.asarray(D_grid, dtype=np.float64)
    dD_min = (D.max() - D.min()) / 64 if dD_min is None else dD_min
    timer = Timer()
    pts = lambda Ds: [(float(D), rho, v0, n) for D in Ds]
    m, e = _converge(pts(D), tol, chunk, sample_every, max_steps, seed, timer)
    log = [("grid", len(D))]
    D, P, E, rounds = _refine(D, m, e, pts, tol, gap, dD_min, chunk, sample_every, max_steps, seed, timer, 1, log,
                              max_rounds)
    return Curve(D, P, E, timer.seconds, timer.particle_steps, len(D), rounds, log, timer.device_seconds)


def steered_points(s: Surrogate, D_lo, D_hi, rho, v0, n, count, gap, depth=3):
    """A uniform grid of `count` noise values plus the midpoints the predicted curve says the
    simulated refinement would add: wherever neighbouring predicted values differ by more than
    `gap`, the midpoint is added, up to `depth` times.  The refinement that a uniform grid needs a
    second (and third) sequential round for is simulated in the first round."""
    D = np.linspace(D_lo, D_hi, count)
    for _ in range(depth):
        P, _, _ = predicted_curve(s, D, rho, v0, n)
        bad = np.nonzero(np.abs(np.diff(P)) > gap)[0]
        if bad.size == 0:
            break
        D = np.sort(np.concatenate([D, 0.5 * (D[bad] + D[bad + 1])]))
    return D


def replicas_for(s: Surrogate, D, rho, v0, n, tol, max_reps=16, per_rep=8.0):
    """Replicas per point from the predicted polarisation variance: var / (per_rep tol^2), at least
    one -- the noisy points near the transition run as several shorter independent runs side by
    side in the batch instead of one long run."""
    _, var, _ = predicted_curve(s, np.asarray(D, dtype=np.float64), rho, v0, n)
    return np.clip(np.ceil(var / (per_rep * tol * tol)), 1, max_reps).astype(int)


def grid_size(s: Surrogate, D_lo, D_hi, rho, v0, n, gap):
    """Grid points for the predicted total variation of P plus the end points."""
    Dg = np.linspace(D_lo, D_hi, 512)
    Pm, _, _ = predicted_curve(s, Dg, rho, v0, n)
    return int(np.ceil(np.abs(np.diff(Pm)).sum() / gap)) + 3


def steered_curve(s: Surrogate, D_lo, D_hi, rho, v0, n, count=None, tol=0.02, gap=None, dD_min=None, chunk=2000,
                  sample_every=10, max_steps=200000, seed=1, max_rounds=6, max_reps=16):
    """Surrogate-steered points (steered_points: a grid of `count` plus its predicted refinement) in
    one batched engine, each with as many replicas as its predicted variance asks for, then
    simulated refinement where still needed."""
    gap = 3 * tol if gap is None else gap
    dD_min = (D_hi - D_lo) / 64 if dD_min is None else dD_min
    timer = Timer()
    if count is None:
        count = grid_size(s, D_lo, D_hi, rho, v0, n, gap)
    D = steered_points(s, D_lo, D_hi, rho, v0, n, count, gap)
    pts = lambda Ds: [(float(x), rho, v0, n) for x in Ds]
    reps_fn = lambda Ds: replicas_for(s, Ds, rho, v0, n, tol, max_reps)
    reps = reps_fn(D)
    m, e = _converge(pts(D), tol, chunk, sample_every, max_steps, seed, timer, reps=reps)
    log = [("steered", count, len(D), int(reps.sum()))]
    D, P, E, rounds = _refine(D, m, e, pts, tol, gap, dD_min, chunk, sample_every, max_steps, seed, timer, 1, log,
                              max_rounds, reps_fn)
    return Curve(D, P, E, timer.seconds, timer.particle_steps, len(D), rounds, log, timer.device_seconds)

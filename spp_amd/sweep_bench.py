"""Second headline measurement: time to a polarisation-versus-noise curve resolved to +-tol.

    python -m spp_amd.sweep_bench [--n 1024] [--tol 0.02] [--train-points 2048] ...

Prints one JSON line.  Workload: continuous-time Vicsek (unit radius, g = 1, RK2, dt = 0.01) at
density 1 and speed 0.5, noise D over [0, 1.5]; error bars by the reference's autocorrelation
estimator (polarization-analysis.py:29-62).

* plain   -- one engine per point of a uniform grid, points one after another, each run until its
             error bar is <= tol (the reference's independent runs; `--plain-max-steps` caps a
             point), grid refined where neighbours differ by more than 3 tol (down to spacing 1.5/64);
* uniform -- the same uniform grid as systems of one batched engine, refined the same way: the
             batching without the network;
* steered -- the surrogate network trained on simulator output predicts the curve; its points go
             along the predicted arc length (with replicas where the predicted variance is high)
             and run together in one batched engine, refined the same way by simulation.
Times are device seconds of the engine calls (the engine's HIP events on its stream around each
stepping call, spp_info.device_ms), with the wall-clock seconds beside them.  `value` is the
steered sweep's device time; `speedup_vs_plain` compares it with the plain sweep,
`network_gain` with the uniform batched sweep.  The steered figure is also reported with the cost
of producing the training set and training the network (reusable across curves).  The network's
own prediction is judged against the plain runs' points, which it never saw
(`surrogate_rms_error_vs_plain`).
`--active-rounds` adds rounds of 128 points of highest predicted variance near the transition,
simulated and added to the training set before the curve is drawn.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from spp_amd import sweep as SW  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--rho", type=float, default=1.0)
    ap.add_argument("--v0", type=float, default=0.5)
    ap.add_argument("--tol", type=float, default=0.02)
    ap.add_argument("--grid", type=int, default=0, help="plain grid points (0: as many as the steered start)")
    ap.add_argument("--train-points", type=int, default=2048)
    ap.add_argument("--train-sizes", default="256,512,1024")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--chunk", type=int, default=4000)
    ap.add_argument("--max-steps", type=int, default=200000, help="step cap per point of the batched sweeps")
    ap.add_argument("--plain-max-steps", type=int, default=2000000, help="step cap per point of the plain sweep")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--active-rounds", type=int, default=0,
                    help="variance-steered rounds of 128 simulated points added to the training set")
    a = ap.parse_args(argv)

    sizes = tuple(int(s) for s in a.train_sizes.split(","))
    t0 = time.perf_counter()
    data = SW.make_dataset(a.train_points, sizes=sizes, seed=a.seed, D_range=(0.0, 1.5), rho_range=(0.25, 4.0),
                           v0_range=(0.25, 1.0), per_batch=64)
    t_data = time.perf_counter() - t0
    print(f"dataset: {len(data.y)} points in {t_data:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    s = SW.train_surrogate(data, epochs=a.epochs, batch=256, seed=a.seed)
    for r in range(a.active_rounds):
        s, data, _, _, _ = SW.active_round(s, data, per_round=128, seed=a.seed + 101 * (r + 1), sizes=sizes,
                                           epochs=a.epochs)
    t_train = time.perf_counter() - t0
    print(f"training: {t_train:.1f} s ({s.train_ms:.0f} ms on the device)", file=sys.stderr, flush=True)
    kw = dict(tol=a.tol, chunk=a.chunk, sample_every=20, max_steps=a.max_steps, seed=a.seed + 100)
    res = lambda c: c.resolved(a.tol, 3 * a.tol, 1.5 / 64)
    summary = lambda c: {"device_seconds": c.device_seconds, "wall_seconds": c.seconds, "points": c.points_simulated,
                         "rounds": c.rounds, "particle_steps": c.particle_steps, "resolved": res(c),
                         "max_err": float(c.err.max())}
    st = SW.steered_curve(s, 0.0, 1.5, a.rho, a.v0, a.n, **kw)
    print(f"steered: {st.points_simulated} points, {st.device_seconds:.2f} device s, log {st.log}", file=sys.stderr,
          flush=True)
    grid = a.grid or st.log[0][1]
    un = SW.batched_curve(np.linspace(0.0, 1.5, grid), a.rho, a.v0, a.n, **kw)
    print(f"uniform: {un.points_simulated} points, {un.device_seconds:.2f} device s, log {un.log}", file=sys.stderr,
          flush=True)
    out = {
        "metric": "device time to a polarisation-vs-noise curve at +-%g" % a.tol,
        "unit": "s",
        "higher_is_better": False,
        "workload": "continuous-time Vicsek N=%d, density %g, v0 %g, D in [0, 1.5], RK2 dt 0.01" % (a.n, a.rho, a.v0),
        "value": st.device_seconds,
        "steered": summary(st),
        "uniform_batched": summary(un),
        "network_gain": un.device_seconds / st.device_seconds,
        "surrogate_cost": {"dataset_seconds": t_data, "dataset_points": int(len(data.y)), "train_seconds": t_train,
                           "train_device_ms": s.train_ms},
    }
    if not a.skip_plain:
        pkw = dict(kw, max_steps=a.plain_max_steps)
        pl = SW.plain_curve(np.linspace(0.0, 1.5, grid), a.rho, a.v0, a.n, **pkw)
        D, P, E = pl.D, pl.P, pl.err
        # the plain sweep refines its uniform grid by simulation, one point at a time
        rounds = 1
        dev_s, wall_s, steps = pl.device_seconds, pl.seconds, pl.particle_steps
        while rounds < 6:
            o = np.argsort(D)
            D, P, E = D[o], P[o], E[o]
            bad = np.nonzero((np.abs(np.diff(P)) > 3 * a.tol) & (np.diff(D) > 2 * 1.5 / 64))[0]
            if bad.size == 0:
                break
            mids = 0.5 * (D[bad] + D[bad + 1])
            add = SW.plain_curve(mids, a.rho, a.v0, a.n, **dict(pkw, seed=kw["seed"] + 17 * rounds))
            D, P, E = np.concatenate([D, add.D]), np.concatenate([P, add.P]), np.concatenate([E, add.err])
            dev_s += add.device_seconds
            wall_s += add.seconds
            steps += add.particle_steps
            rounds += 1
        o = np.argsort(D)
        D, P, E = D[o], P[o], E[o]
        ok_gap = (np.abs(np.diff(P)) <= 3 * a.tol) | (np.diff(D) <= 2 * 1.5 / 64)
        print(f"plain: {D.size} points, {dev_s:.2f} device s", file=sys.stderr, flush=True)
        out["plain"] = {"device_seconds": dev_s, "wall_seconds": wall_s, "points": int(D.size), "rounds": rounds,
                        "particle_steps": steps, "resolved": bool(np.all(E <= a.tol) and np.all(ok_gap)),
                        "max_err": float(E.max())}
        # the network judged against the held-out plain runs (it never saw these points)
        pred, _, _ = SW.predicted_curve(s, D, a.rho, a.v0, a.n)
        out["surrogate_rms_error_vs_plain"] = float(np.sqrt(np.mean((pred - P) ** 2)))
        out["speedup_vs_plain"] = dev_s / st.device_seconds
        out["speedup_vs_plain_incl_training"] = dev_s / (st.device_seconds + t_data + t_train)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

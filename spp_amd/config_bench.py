"""Throughput of the engine on four more workloads (one JSON line each).

    python -m spp_amd.config_bench [--only vicsek4096,vesicle,batch256,lyapunov64] [--scale 1.0]

* vicsek4096 -- continuous-time Vicsek, N = 4096 in a box of side 64, R = 1, v0 = 0.5, g = 1, D = 0.1,
               RK2 dt = 0.01, Verlet radius 1.3 (displacement rebuilds);
* vesicle    -- vesicle: a ring of 1024 beads at radius 180 (bonds k = 100, rest length 1, bending 10)
               around 8192 active Brownian swimmers placed uniformly inside (speed 1: F0 = 2 at mobility
               0.5, D = 0.1), harmonic repulsion k = 10 over radius 1 between all, RK2 dt = 1e-3;
* batch256   -- 256 independent Vicsek boxes in one batched engine, N log-uniform in [1024, 65536],
               density U[0.25, 4], D U[0, 0.5], v0 = 0.5, RK2 dt = 0.01, run in calls of 100 steps;
* lyapunov64 -- 64 Lyapunov replicas of a repulsive Vicsek box (N = 16,384, density 1, the harmonic
               repulsion of the N = 262,144 workload; triangular-lattice start, random headings), each
               with 4 perturbed clones, RK4 dt = 0.005, tangent vectors renormalised every 10 steps;
               the autotuner first picks the slot kernel's depth (4 or 8 records in flight).

Each line gives particle-steps/s over the timed steps (inputs resident on the device), the
wall-clock per step, and the fused stage-1 kernel's duration (HIP events around back-to-back
launches, after the timed steps) with its algorithmic bytes over that duration as a fraction of
the 8 TB/s HBM peak (the ring's bond and bending gathers of the vesicle's 1024 beads are not
counted).  `--scale` shortens the timed step counts.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import spp_amd as S  # noqa: E402
from spp_amd.autotune import autotune  # noqa: E402

HBM_PEAK_GBS = 8000.0


def stage1_bytes(kbar, rk4=False):
    """Algorithmic bytes per particle of the fused stage 1 (the minimum it must move, noise drawn in
    registers, a neighbour's record counted once, as its owner's): read x, y, theta (24) and the
    list (4 kbar), write the midpoint x, y, theta (24); RK4's stage 1 also writes the accumulator
    (fx, fy, tau: 24)."""
    return 48.0 + 4.0 * kbar + (24.0 if rk4 else 0.0)


def timed(e, n, steps, chunk, warm):
    e.multi_step(warm)
    e.synchronize()
    i0 = e.info()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        e.multi_step(k)
        done += k
    e.synchronize()
    el = time.perf_counter() - t0
    i1 = e.info()
    return n * steps / el, el / steps * 1e3, i1.rebuilds - i0.rebuilds


def report(name, workload, n, steps, rate, ms, rebuilds, e, rk4=False, kernel_variant=S.KERNEL_SLOT):
    us = e.time_stage(1, 50)
    b = stage1_bytes(e.info().mean_neighbors, rk4) * n
    out = {"config": name, "workload": workload, "particles": n, "steps": steps, "particle_steps_per_s": rate,
                 "ms_per_step": ms, "rebuilds": rebuilds, "stage1_us": us, "mean_full_list": e.info().mean_neighbors,
                 "stage1_bytes": b, "stage1_hbm_frac": b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                 "kernel_variant": kernel_variant}
    print(json.dumps(out), flush=True)


def vicsek4096(scale):
    p = S.default_params(S.MODEL_CONTINUOUS).update(
        n=4096, Lx=32.0, Ly=32.0, dt=0.01, v0=0.5, g=1.0, R_align=1.0, alpha=0.0, D_rot=0.1, integrator=S.RK2,
        r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT, seed=1)
    e = S.Engine(p)
    e.set_state(*S.random_state(1, p.n, p.Lx, p.Ly))
    steps = int(2000 * scale)
    rate, ms, rb = timed(e, p.n, steps, steps, 100)
    report("vicsek4096", "continuous Vicsek N=4096, side 64, RK2 dt 0.01", p.n, steps, rate, ms, rb, e)


def vesicle(scale):
    nm, ns, R = 1024, 8192, 180.0
    p = S.default_params(S.MODEL_ABP).update(
        n=nm + ns, n_membrane=nm, Lx=R + 6, Ly=R + 6, integrator=S.RK2, dt=1e-3, repulsion=S.REP_HARMONIC,
        A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, r_list=1.3, D_rot=0.1, v0=2.0, mob_par=0.5, mob_perp=0.5,
        mob_rot=1.0, k_bond=100.0, l_bond=1.0, kappa_bend=10.0, seed=6)
    mx, my = S.ring_membrane(nm, 1.0, R)
    sx, sy, st = S.disk_state(11, ns, R - 1.0)
    e = S.Engine(p)
    e.set_state(np.concatenate([mx, sx]), np.concatenate([my, sy]), np.concatenate([np.zeros(nm), st]))
    steps = int(2000 * scale)
    rate, ms, rb = timed(e, p.n, steps, steps, 100)
    report("vesicle", "vesicle: 1024-bead ring (R 180) + 8192 swimmers, RK2 dt 1e-3", p.n, steps, rate, ms, rb, e)


def batch256(scale):
    rng = np.random.default_rng(4)
    systems, xs = [], []
    for k in range(256):
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(65536)))))
        rho = rng.uniform(0.25, 4.0)
        L = 0.5 * math.sqrt(n / rho)
        systems.append(S.system(n, L, v0=0.5, g=1.0, D_rot=rng.uniform(0, 0.5), seed=100 + k))
        xs.append(S.random_state(200 + k, n, L, L))
    p = S.default_params(S.MODEL_CONTINUOUS).update(dt=0.01, R_align=1.0, alpha=0.0, integrator=S.RK2, r_list=1.3,
                                                         rebuild=S.REBUILD_DISPLACEMENT)
    e = S.Engine(p, systems)
    e.set_state(*[np.concatenate([v[i] for v in xs]) for i in range(3)])
    n = sum(s.n for s in systems)
    steps = int(500 * scale)
    rate, ms, rb = timed(e, n, steps, 100, 100)
    report("batch256", "256 ragged Vicsek boxes (N 1k-64k, density 0.25-4) in one engine, RK2 dt 0.01, 100-step calls",
                 n, steps, rate, ms, rb, e)


def lyapunov64(scale):
    R, K, n = 64, 4, 16384
    L = 0.5 * math.sqrt(n)
    p = S.default_params(S.MODEL_REPULSIVE).update(
        dt=0.005, R_align=1.0, repulsion=S.REP_HARMONIC, A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, integrator=S.RK4,
        r_list=1.3, rebuild=S.REBUILD_DISPLACEMENT)
    systems = [S.system(n, L, v0=1.0, g=1.0, D_rot=0.05, seed=20 + r) for r in range(R) for _ in range(K + 1)]
    # uniform random positions and headings per replica: somewhere among 5 million particles the
    # first steps push deep overlaps apart faster than the half skin per step, which the engine
    # handles on its wide list (spp_info.wide_rebuilds) until the particles slow down
    xs, ys, ts = [], [], []
    for r in range(R):
        x, y, t = S.random_state(50 + r, n, L, L)
        for _ in range(K + 1):
            xs.append(x)
            ys.append(y)
            ts.append(t)
    x, y, th = np.concatenate(xs), np.concatenate(ys), np.concatenate(ts)
    # the autotuner picks four or eight neighbour records in flight per lane by timing plain steps
    # of this batch (about a second); the winner is reported with the line
    tuned = autotune(p, x, y, th, systems=systems, steps=100, warm=20, variants=[S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP],
                     list_factors=(1.3,), cell_scales=(1.0,))
    p = tuned.params
    e = S.Engine(p, tuned.systems)
    e.set_state(x, y, th)
    e.lyapunov_init(R, K, 1e-7, rescale=True, seed=4)
    e.lyapunov_run(50, 10)
    e.synchronize()
    steps = int(200 * scale)
    N = R * (K + 1) * n
    t0 = time.perf_counter()
    e.lyapunov_run(steps, 10)
    e.synchronize()
    el = time.perf_counter() - t0
    report("lyapunov64", "64 Lyapunov replicas x (1 + 4 clones) x 16384, RK4 dt 0.005, renormalised every 10 steps", N,
                 steps, N * steps / el, el / steps * 1e3, e.info().rebuilds, e, rk4=True,
           kernel_variant=p.kernel_variant)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="vicsek4096,vesicle,batch256,lyapunov64")
    ap.add_argument("--scale", type=float, default=1.0)
    a = ap.parse_args(argv)
    fns = {"vicsek4096": vicsek4096, "vesicle": vesicle, "batch256": batch256, "lyapunov64": lyapunov64}
    for name in a.only.split(","):
        fns[name](a.scale)


if __name__ == "__main__":
    main()

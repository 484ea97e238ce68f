"""Benchmark of the per-time-step particle engine on one MI355X.

Workload: repulsive Vicsek at N = 262,144 in a periodic box of side 512 (density 1), v0 = 1,
alignment radius 1, harmonic repulsion k = 10 over radius 1, angular noise D = 0.05, RK2 with
dt = 0.005 in FP64, Verlet list radius 1.3 rebuilt on displacement.  A step is one full RK2 step
(two fused interact+move stages, plus the amortised neighbour-list rebuilds and checks).

Prints one JSON line: particle-steps/s over the timed steps, the roofline of the dominant kernel
(algorithmic bytes per launch over its HIP-event-timed average duration on the engine stream) and
the CPU baseline (concurrent copies of the C oracle on the host cores, and one copy alone, each
on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def workload_params(S, n, L, D, variant=0):
    p = S.default_params(S.MODEL_REPULSIVE)
    p.update(n=n, Lx=L, Ly=L, dt=0.005, v0=1.0, g=1.0, R_align=1.0, repulsion=S.REP_HARMONIC,
             A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, D_rot=D, integrator=S.RK2, r_list=1.3,
             rebuild=S.REBUILD_DISPLACEMENT, seed=2, kernel_variant=variant)
    return p


def stage_bytes(kbar):
    """Algorithmic bytes per particle of each fused RK2 stage: the minimum a stage must move, FP64
    SoA, noise generated in registers, a neighbour's record counted once (as its owner's), kbar the
    mean full-list length.
      stage 1: read x,y,theta (24) + list (4 kbar); write midpoint x,y,theta (24)    = 48 + 4 kbar
      stage 2: read midpoint (24) + x_u,y_u,theta_old (24) + list (4 kbar);
               write x_w,y_w,x_u,y_u,theta (40)                                      = 88 + 4 kbar
    A whole RK2 step is 136 + 8 kbar, plus about 100 B per rebuild."""
    return 48 + 4 * kbar, 88 + 4 * kbar


def stage_impl_bytes(kbar):
    """Bytes the engine's fused RK2 stages actually move per particle (see stage_bytes for the
    algorithmic minimum; a neighbour's record is counted once, as its owner's):
      stage 1: read x,y (16) + theta (8) + list length (4) + list (4 kbar) + id (4, noise counter)
               + build position (8, displacement check); write midpoint x,y (16) + theta (8)
               + noise (8)                                                          = 72 + 4 kbar
      stage 2: read midpoint x,y,theta (24) + length (4) + list (4 kbar) + noise (8) + theta_old
               (8) + x_u,y_u (16) + build position (8); write x,y (16) + theta (8) + x_u,y_u (16)
                                                                                    = 108 + 4 kbar"""
    return 72 + 4 * kbar, 108 + 4 * kbar


REBUILD_BYTES = 100.0  # algorithmic bytes per particle and rebuild (binning, sort, list)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_rk2_c2_n262144_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter on a copy kernel of known
    size) and a label naming the summary and the kernel build it was measured on; (None, None)
    when the summary does not cover this kernel."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
        t = d["kernels"][kernel]
        return t["hbm_bytes_per_launch"], "%s (%s)" % (os.path.relpath(TRAFFIC_FILE, ROOT), d["measured_on"])
    except (OSError, KeyError, ValueError):
        return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(p, x, y, th, budget_s, threads=1, library=None):
    """Particle-steps/s of `threads` concurrent copies of the C oracle on the same workload (the
    reference's replica pattern, serial/mpi-main.cpp:157-160), each stepped for about budget_s
    seconds after 2 warm-up steps.  ctypes releases the GIL inside the C calls, so the copies run
    on separate cores.  Returns (aggregate rate, per-copy rates, total steps)."""
    import threading

    import oracle

    copies = [oracle.Oracle(p, library=library) for _ in range(threads)]
    rates = [0.0] * threads
    steps = [0] * threads

    def run(k):
        o = copies[k]
        o.set_state(x, y, th)
        o.multi_step(2)  # warm-up (includes a list build)
        t0 = time.perf_counter()
        while True:
            o.multi_step(5)
            steps[k] += 5
            el = time.perf_counter() - t0
            if el >= budget_s:
                break
        rates[k] = p.n * steps[k] / el

    pool = [threading.Thread(target=run, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    return sum(rates), rates, sum(steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--noise", type=float, default=0.05)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16, help="concurrent oracle copies (host cores) for the CPU leg")
    ap.add_argument("--variant", type=int, default=0, help="interaction-kernel variant (spp_params.kernel_variant)")
    ap.add_argument("--time-reps", type=int, default=50, help="launches per stage-kernel timing")
    ap.add_argument("--stage-samples", type=int, default=10, help="stage timings along the replayed timed region")
    ap.add_argument("--calibration-copy", type=int, default=0,
                    help="profiling runs: also snapshot an engine of this many particles twice (a plain copy of "
                         "44 B per particle read and written, far larger than the L2) to calibrate byte counters")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs one box on one GPU; multi-GPU runs are not part of this "
              "build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import spp_amd as S

    n = args.n
    L = 0.5 * float(np.sqrt(n))  # density 1: side 2L = sqrt(N)
    p = workload_params(S, n, L, args.noise, args.variant)
    x, y, th = S.random_state(7, n, L, L)
    e = S.Engine(p)
    e.set_state(x, y, th)
    e.multi_step(args.warmup)
    e.synchronize()

    e.snapshot(98)  # the timed region's starting state, for the per-kernel replay below
    e.synchronize()
    info0 = e.info()
    t0 = time.perf_counter()
    e.multi_step(args.steps)
    e.synchronize()
    t1 = time.perf_counter()
    info1 = e.info()
    el = t1 - t0
    value = n * args.steps / el
    ms_per_step = el / args.steps * 1e3

    # Per-kernel HIP-event timing over a separate profiled run (the timed run above is not
    # perturbed by the events).
    e.profile(True)
    e.multi_step(args.profile_steps)
    prof = e.profile_read()
    e.profile(False)
    info2 = e.info()
    kbar = info2.mean_neighbors
    b1, b2 = stage_bytes(kbar)
    i1, i2 = stage_impl_bytes(kbar)
    # mean steps between rebuilds over the timed and the profiled runs together
    nreb = (info1.rebuilds - info0.rebuilds) + (info2.rebuilds - info1.rebuilds)
    interval = (args.steps + args.profile_steps) / max(1, nreb)
    # A plain copy of known size (the snapshot copy: 44 B per particle read and written) -- the
    # calibration kernel for hardware byte counters in profiling runs.
    e.snapshot(99)
    e.synchronize()
    if args.calibration_copy > 0:
        nc = args.calibration_copy
        Lc = 0.5 * float(np.sqrt(nc))
        ec = S.Engine(workload_params(S, nc, Lc, args.noise))
        ec.set_state(*S.random_state(11, nc, Lc, Lc))
        ec.snapshot(1)
        ec.snapshot(2)
        ec.synchronize()
        del ec
    # Dominant kernel, measured along the timed region's trajectory: the region is replayed from its
    # starting state and, every steps/K steps, each fused RK2 stage is timed over back-to-back
    # launches between two HIP events on the engine's stream (outputs to scratch, state untouched).
    # The timed run itself carries no events.
    end_us = {"stage_rk2_1": e.time_stage(1, args.time_reps), "stage_rk2_2": e.time_stage(2, args.time_reps)}
    e.restore(98)
    K = max(1, args.stage_samples)
    samples = []
    for _ in range(K):
        e.multi_step(max(1, args.steps // K))
        kb = e.info().mean_neighbors
        samples.append((e.time_stage(1, args.time_reps), e.time_stage(2, args.time_reps), stage_bytes(kb)))
    stage_us = {"stage_rk2_1": float(np.mean([q[0] for q in samples])),
                "stage_rk2_2": float(np.mean([q[1] for q in samples]))}
    dom = max(stage_us, key=stage_us.get)
    col = 0 if dom.endswith("_1") else 1
    dom_ms = stage_us[dom] * 1e-3
    dom_bytes = float(np.mean([q[2][col] for q in samples])) * n
    achieved = sum(q[2][col] * n for q in samples) / (sum(q[col] for q in samples) * 1e-6) / 1e9
    impl_bytes = (i1, i2)[col] * n
    kernel_ms = {k: round(v[0] * 1e3, 3) for k, v in prof.items()}  # microseconds
    # whole step against the HBM roofline: 136 + 8 kbar per particle-step plus the rebuilds
    step_bytes = b1 + b2 + REBUILD_BYTES / interval
    traffic, traffic_src = measured_traffic(dom) if args.n == 262144 and args.variant == 0 else (None, None)
    out = {
        "metric": "particle-steps/s at N=262k RK2",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform random positions and headings (seed 7), project Philox noise",
        "config": {
            "workload": "repulsive Vicsek N=%d, side %.0f (density 1), v0=1, R_a=1, harmonic k=10 rc=1, "
                        "D=%g, RK2 dt=0.005, Verlet r_list=1.3 rebuilt on displacement" % (n, 2 * L, args.noise),
            "n": n,
            "box_side": 2 * L,
            "integrator": "rk2",
            "kernel_variant": args.variant,
            "mean_full_list": round(kbar, 3),
            "rebuilds_in_timed": info1.rebuilds - info0.rebuilds,
            "host_syncs_in_timed": info1.host_syncs - info0.host_syncs,
            "rollbacks_in_timed": info1.rollbacks - info0.rollbacks,
            "steps_per_rebuild": round(interval, 2),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "bytes_per_launch": dom_bytes,
            "bytes_rule": "48 + 4 kbar (stage 1) / 88 + 4 kbar (stage 2) per particle, kbar = mean full-list length",
            "implementation_bytes_per_launch": impl_bytes,
            "implementation_frac": impl_bytes / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "avg_launch_us": dom_ms * 1e3,
        },
        "step_roofline": {
            "bytes_per_particle_step": step_bytes,
            "bytes_rule": "136 + 8 kbar + 100 / steps_per_rebuild",
            "achieved": step_bytes * value / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_bytes * value / 1e9 / HBM_PEAK_GBS,
        },
        "stage_kernel_us": {k: round(v, 3) for k, v in stage_us.items()},
        "stage_kernel_us_at_end": {k: round(v, 3) for k, v in end_us.items()},
        "kernel_avg_us_event_bracketed": kernel_ms,
    }

    if not args.no_cpu_baseline:
        # Bounded samples of the same workload on the host: the C oracle (AoS, per-cell lists,
        # Newton-3 half Verlet list -- the reference's data structures), first on one core, then
        # as one copy per core of the host share, all at once.
        # The oracle is compiled here for this host (-O3 -march=native, the flags the CPU-baseline
        # plan names); the portable -O3 build of the tests is the fallback.
        import oracle

        try:
            lib_path, flags = oracle.build_native(), " ".join(oracle.NATIVE_FLAGS)
        except Exception:  # no compiler on this host
            lib_path, flags = oracle.build(), "-O3 (portable build; native build unavailable)"
        library = oracle.load_library(lib_path)
        cp = workload_params(S, n, L, args.noise)
        one, _, steps1 = cpu_baseline(cp, x, y, th, args.cpu_seconds / 2, 1, library)
        P = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        agg, rates, stepsP = cpu_baseline(cp, x, y, th, args.cpu_seconds / 2, P, library)
        out["cpu_baseline"] = {
            "value": agg,
            "unit": "particle-steps/s",
            "cores": P,
            "kind": "port",
            "sample": "same workload (N=%d, RK2, displacement rebuilds, same start state) on %s (%d logical CPUs "
                      "visible): %d concurrent copies of the C oracle (AoS particles, per-cell lists, Newton-3 half "
                      "Verlet list), %s %s, %d steps in %.1f s each after 2 warm-up steps (aggregate); one copy "
                      "alone: %.4g particle-steps/s over %d steps, x%d = %.4g" % (
                          n, cpu_model(), os.cpu_count() or 0, P, oracle.compiler_version(), flags, stepsP // P,
                          args.cpu_seconds / 2, one, steps1, P, one * P),
            "single_core": one,
            "gpu_over_cpu": value / agg,
            "gpu_over_ideal_cores": value / (one * P),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner for the engine's launch-shape knobs, by timing short runs of the real engine.

The knobs do not change the physics; they trade gather strategy against list length and rebuild
frequency:

* ``kernel_variant`` -- the interaction tile strategy: one lane per slot with L2 gathers (0), one
  workgroup per system running the whole call in one launch (3; systems of at most 4096
  particles), or the slot kernel with eight neighbour records in flight per lane at every size (4;
  the slot kernel takes eight only below 393,216 particles);
* ``r_list`` -- the Verlet radius: a wider skin lengthens every list but rebuilds less often;
* the cell size -- ``ncell_x/y`` below floor(2L / r_list) makes cells wider than the list radius
  (more candidates per list build, fewer empty cells).

Every candidate runs ``warm`` steps and is then timed over ``steps`` steps from the same state
(the reference's fixed ``cell_update_period`` and cell divisor, parameters.h:80-99, were chosen by
hand for one CPU; here they are measured per workload and device).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

from . import (KERNEL_PERSIST, KERNEL_SLOT, KERNEL_SLOT_DEEP, MODEL_ABP, MODEL_REPULSIVE, PERSIST_MAX_N, REP_NONE,
               Engine, SppError, System)


def interaction_cutoff(p):
    """The largest interaction radius (mirrors the engine's rule)."""
    if p.model == MODEL_ABP:
        return p.rc_rep if p.repulsion != REP_NONE else 0.0
    c = p.R_align
    if p.model == MODEL_REPULSIVE and p.repulsion != REP_NONE:
        c = max(c, p.rc_rep)
    return c


@dataclass
class Trial:
    kernel_variant: int
    r_list: float
    cell_scale: float
    rate: float                       # particle-steps per second (0 if the candidate is invalid)
    error: str = ""


@dataclass
class Tuned:
    params: object
    systems: list | None
    trials: list = field(default_factory=list)

    @property
    def best(self) -> Trial:
        return max(self.trials, key=lambda t: t.rate)


def _with_cells(systems, r_list, scale):
    out = []
    for s in systems:
        q = System()
        for f, _ in System._fields_:
            setattr(q, f, getattr(s, f))
        q.ncell_x = max(3, int(math.floor(2 * s.Lx / (r_list * scale))))
        q.ncell_y = max(3, int(math.floor(2 * s.Ly / (r_list * scale))))
        out.append(q)
    return out


def autotune(params, x, y, theta, systems=None, steps=200, warm=50, variants=None, list_factors=(1.15, 1.3, 1.5),
             cell_scales=(1.0, 1.5)) -> Tuned:
    """Time every candidate on the state (x, y, theta) and return the fastest configuration."""
    cut = interaction_cutoff(params)
    if variants is None:
        variants = [KERNEL_SLOT]
        if params.model != MODEL_ABP:
            variants.append(KERNEL_SLOT_DEEP)
        sizes = [s.n for s in systems] if systems is not None else [params.n]
        if params.model != MODEL_ABP and max(sizes) <= PERSIST_MAX_N:
            variants.append(KERNEL_PERSIST)
    n = int(len(x))
    trials = []
    best = None
    for v in variants:
        for f in list_factors:
            r_list = cut * f if cut > 0 else params.r_list
            for scale in cell_scales:
                p = params.copy().update(kernel_variant=v, r_list=r_list)
                syst = None
                if systems is None:
                    p.ncell_x = max(3, int(math.floor(2 * p.Lx / (r_list * scale))))
                    p.ncell_y = max(3, int(math.floor(2 * p.Ly / (r_list * scale))))
                else:
                    syst = _with_cells(systems, r_list, scale)
                try:
                    e = Engine(p, syst)
                    e.set_state(x, y, theta)
                    e.multi_step(warm)
                    e.synchronize()
                    t0 = time.perf_counter()
                    e.multi_step(steps)
                    e.synchronize()
                    rate = n * steps / (time.perf_counter() - t0)
                    e.close()
                    trials.append(Trial(v, r_list, scale, rate))
                except SppError as err:  # e.g. cells narrower than the list radius
                    trials.append(Trial(v, r_list, scale, 0.0, str(err)))
                    continue
                if best is None or rate > best[0]:
                    best = (rate, p, syst)
    if best is None:
        raise SppError(-1, "no valid candidate")
    return Tuned(best[1], best[2], trials)

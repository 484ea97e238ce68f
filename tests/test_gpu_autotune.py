"""GPU: the autotuner times real candidates and its choice changes speed, not physics."""
import math

import numpy as np
import pytest

import spp_amd as S
from tests.helpers import angle_diff, c2_params

pytestmark = pytest.mark.gpu


def test_autotune_single_box_keeps_the_physics():
    from spp_amd.autotune import autotune

    n, L = 65536, 128.0
    p = c2_params(n=n, L=L)
    x, y, th = S.random_state(4, n, L, L)
    tuned = autotune(p, x, y, th, steps=60, warm=20, cell_scales=(1.0, 1.5))
    # slot and deep slot variants x three list radii x two cell sizes
    assert len(tuned.trials) == 2 * 3 * 2
    assert {t.kernel_variant for t in tuned.trials} == {S.KERNEL_SLOT, S.KERNEL_SLOT_DEEP}
    # a skin too thin for the workload's fastest particles is an invalid candidate, with its reason
    for t in tuned.trials:
        assert (t.rate > 0) != bool(t.error), t
    assert any(t.rate > 0 and t.r_list == pytest.approx(1.3) for t in tuned.trials)
    assert tuned.best.rate == max(t.rate for t in tuned.trials) > 0
    a, b = S.Engine(p), S.Engine(tuned.params)
    for e in (a, b):
        e.set_state(x, y, th)
        e.multi_step(50)
    sa, sb = a.get_state(), b.get_state()
    assert np.abs(sa["xu"] - sb["xu"]).max() < 1e-9
    assert angle_diff(sa["theta"], sb["theta"]).max() < 1e-9


def test_autotune_batch_sets_per_system_cells():
    from spp_amd.autotune import autotune

    p = S.default_params(S.MODEL_CONTINUOUS).update(dt=0.01, R_align=1.0, integrator=S.RK2, r_list=1.3)
    rng = np.random.default_rng(1)
    systems = [S.system(n, 0.5 * math.sqrt(n), v0=0.5, g=1.0, D_rot=0.2, seed=k) for k, n in
               enumerate(rng.integers(500, 4000, 6))]
    xs = [S.random_state(k, s.n, s.Lx, s.Ly) for k, s in enumerate(systems)]
    x, y, th = (np.concatenate([v[i] for v in xs]) for i in range(3))
    tuned = autotune(p, x, y, th, systems=systems, steps=60, warm=20)
    assert tuned.systems is not None and all(s.ncell_x >= 3 for s in tuned.systems)
    a, b = S.Engine(p, systems), S.Engine(tuned.params, tuned.systems)
    for e in (a, b):
        e.set_state(x, y, th)
        e.multi_step(50)
    assert np.abs(a.get_state()["xu"] - b.get_state()["xu"]).max() < 1e-9

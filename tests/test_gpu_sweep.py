"""GPU: the parameter-sweep driver -- batched runs, the training set, the surrogate fitted to it,
and the curves with and without the network (systems of 256 particles, so the whole file runs in
about a minute)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_points_are_deterministic_and_physical():
    from spp_amd import sweep as SW

    pts = [(0.02, 1.0, 0.5, 256), (2.0, 1.0, 0.5, 256), (0.02, 1.0, 0.5, 256)]
    s, final = SW.run_points(pts, 12000, 40, seed=3, ordered=False)
    assert s.shape == (300, 3)
    assert np.all((s >= 0) & (s <= 1))
    # identical parameters but different noise keys: different paths; ordered vs disordered phase
    assert not np.array_equal(s[:, 0], s[:, 2])
    tail = s[150:]
    assert tail[:, 0].mean() > 0.8 and tail[:, 1].mean() < 0.2
    s2, _ = SW.run_points(pts, 12000, 40, seed=3, ordered=False)
    assert np.array_equal(s, s2)


def test_dataset_and_surrogate():
    from spp_amd import sweep as SW

    data = SW.make_dataset(96, sizes=(256,), warm=4000, steps=4000, sample_every=20, seed=5, per_batch=48,
                           D_range=(0.0, 1.5), rho_range=(0.5, 2.0), v0_range=(0.4, 0.6))
    assert data.x.shape == (96, 4) and np.all(np.isfinite(data.y))
    low, high = data.x[:, 0] < 0.2, data.x[:, 0] > 1.2
    assert data.y[low, 0].mean() > data.y[high, 0].mean() + 0.3
    s = SW.train_surrogate(data, epochs=300, batch=48, seed=2)
    pred = s.predict(data.x[:, 0], data.x[:, 1], data.x[:, 2], data.x[:, 3])
    r2 = 1 - np.mean((pred[:, 0] - data.y[:, 0]) ** 2) / np.var(data.y[:, 0])
    assert r2 > 0.8, r2
    assert s.train_ms > 0


def test_steered_sweep_beats_uniform_at_002():
    """The network's gain at the metric's resolution: with the same batched engine, error bars of
    0.02 and the same refinement budget, points placed along the predicted curve (with replicas
    where the predicted variance is high) resolve it in strictly fewer particle-steps and less
    device time than a uniform grid of the same initial size refined by simulation alone.  The
    sweep is deterministic for a given seed."""
    from spp_amd import sweep as SW

    data = SW.make_dataset(192, sizes=(256,), warm=4000, steps=4000, sample_every=20, seed=7, per_batch=96,
                           D_range=(0.0, 1.5), rho_range=(0.5, 2.0), v0_range=(0.4, 0.6))
    s = SW.train_surrogate(data, epochs=300, batch=64, seed=3)
    tol = 0.02
    kw = dict(tol=tol, chunk=4000, sample_every=20, max_steps=400000, seed=11)
    st = SW.steered_curve(s, 0.0, 1.5, 1.0, 0.5, 256, **kw)
    un = SW.batched_curve(np.linspace(0.0, 1.5, st.log[0][1]), 1.0, 0.5, 256, **kw)
    print("steered: %d points, %.3g particle-steps, %.2f device s, log %s; uniform: %d points, %.3g particle-steps, "
          "%.2f device s, resolved %s, log %s" % (st.points_simulated, st.particle_steps, st.device_seconds, st.log,
                                                 un.points_simulated, un.particle_steps, un.device_seconds,
                                                 un.resolved(tol, 3 * tol, 1.5 / 64), un.log))
    assert st.resolved(tol, 3 * tol, 1.5 / 64), st.log
    assert st.particle_steps < un.particle_steps
    assert 0 < st.device_seconds < un.device_seconds
    # an ordered-to-disordered curve
    assert st.P[0] > 0.8 and st.P[-1] < 0.2


def test_variance_steered_round_and_held_out_error():
    """One round of variance-steered sampling: the picked points are the candidates of highest
    predicted variance inside the predicted transition band, every pick is simulated and joins the
    training set, and the retrained network is judged against held-out plain runs (its error is
    measured, and stays a fraction of the curve's span)."""
    from spp_amd import sweep as SW

    data = SW.make_dataset(96, sizes=(256,), warm=3000, steps=3000, sample_every=20, seed=11, per_batch=48,
                           D_range=(0.0, 1.5), rho_range=(0.5, 2.0), v0_range=(0.5, 0.5))
    s = SW.train_surrogate(data, epochs=300, batch=48, seed=4)
    held = SW.simulate_points([(D, 1.0, 0.5, 256) for D in np.linspace(0.1, 1.4, 12)], warm=3000, steps=3000,
                              seed=99)
    before = SW.held_out_error(s, held)
    s2, data2, picks, pv, cv = SW.active_round(s, data, per_round=32, candidates=512, seed=12, sizes=(256,),
                                               rho_range=(0.5, 2.0), v0_range=(0.5, 0.5), warm=3000, steps=3000,
                                               epochs=300, batch=48)
    assert picks.shape == (32, 4) and len(data2.y) == len(data.y) + 32
    assert np.array_equal(data2.x[-32:], picks)  # every pick was simulated and added
    assert np.all(np.isfinite(data2.y[-32:]))
    assert pv.min() >= np.median(cv)  # the most variable candidates were chosen
    after = SW.held_out_error(s2, held)
    print("held-out rms error of P: %.3f before, %.3f after the round" % (before, after))
    assert after < 0.25 and before < 0.25

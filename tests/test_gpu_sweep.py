"""GPU: the parameter-sweep driver -- batched runs, the training set, the surrogate fitted to it,
and the curves with and without the network (small systems, loose error bars, so the whole file
runs in well under a minute)."""
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


def test_steered_sweep_needs_fewer_points_than_uniform():
    """The network's gain: with the same engine and the same error bars, points placed along the
    predicted curve resolve it with fewer simulated points (and particle-steps) than a uniform grid
    of the same initial size refined by simulation alone."""
    from spp_amd import sweep as SW

    data = SW.make_dataset(192, sizes=(256,), warm=4000, steps=4000, sample_every=20, seed=7, per_batch=96,
                           D_range=(0.0, 1.5), rho_range=(0.5, 2.0), v0_range=(0.4, 0.6))
    s = SW.train_surrogate(data, epochs=300, batch=64, seed=3)
    kw = dict(tol=0.06, chunk=4000, sample_every=20, max_steps=40000, seed=11)
    st = SW.steered_curve(s, 0.0, 1.5, 1.0, 0.5, 256, **kw)
    un = SW.batched_curve(np.linspace(0.0, 1.5, st.log[0][1]), 1.0, 0.5, 256, **kw)
    assert st.resolved(0.06, 0.18, 1.5 / 64) and un.resolved(0.06, 0.18, 1.5 / 64), (st.log, un.log)
    assert st.points_simulated <= un.points_simulated, (st.log, un.log)
    # an ordered-to-disordered curve
    assert st.P[0] > 0.8 and st.P[-1] < 0.2

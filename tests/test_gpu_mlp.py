"""GPU: the surrogate network's BF16-MFMA kernels against the CPU oracle (oracle/mlp.py).

Tolerances: the oracle repeats every bfloat16 rounding; what remains is the FP32 summation order of
the matrix products and the last bit of expf, which can flip an isolated bfloat16 rounding (2^-8
relative on one activation).  Norm-wise relative errors are therefore held to 1e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-3


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def make(**kw):
    from spp_amd import surrogate as SG

    return SG.Mlp(**kw)


@pytest.mark.parametrize("shape", [(4, 128, 3, 3), (5, 40, 2, 7), (3, 17, 1, 1)])
@pytest.mark.parametrize("batch", [1, 37, 256, 1000])
def test_forward_matches_oracle(shape, batch):
    import oracle.mlp as M

    n_in, hidden, nh, n_out = shape
    net = make(n_in=n_in, hidden=hidden, n_hidden=nh, n_out=n_out, seed=3)
    rng = np.random.default_rng(batch)
    x = rng.standard_normal((batch, n_in)).astype(np.float32)
    y = net.forward(x)
    want = M.MLPOracle(net.dims, net.get_params()).forward(x)
    assert rel(y, want) < TOL
    assert np.abs(y - want).max() < 1e-2 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("shape", [(4, 128, 3, 3), (5, 40, 2, 7)])
def test_gradients_match_oracle(shape):
    import oracle.mlp as M

    n_in, hidden, nh, n_out = shape
    net = make(n_in=n_in, hidden=hidden, n_hidden=nh, n_out=n_out, seed=5)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((300, n_in)).astype(np.float32)
    t = rng.standard_normal((300, n_out)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n_out).astype(np.float32)
    loss, g = net.loss_grad(x, t, w)
    o = M.MLPOracle(net.dims, net.get_params())
    wl, wg = o.loss_grad(x, t, w)
    assert abs(loss - wl) < 1e-4 * abs(wl)
    assert rel(g, wg) < TOL
    # per layer, so a wrong small layer cannot hide behind a large one
    off = 0
    for din, dout in zip(net.dims[:-1], net.dims[1:]):
        for size in (din * dout, dout):
            assert rel(g[off:off + size], wg[off:off + size]) < 5 * TOL, (din, dout, size)
            off += size


def test_train_step_is_adam_on_the_gradient():
    import oracle.mlp as M

    net = make(seed=9)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((128, 4)).astype(np.float32)
    t = rng.standard_normal((128, 3)).astype(np.float32)
    p0 = net.get_params()
    _, g = net.loss_grad(x, t)
    net.train_step(x, t)
    o = M.MLPOracle(net.dims, p0)
    z = np.zeros_like(p0)
    o.adam_step(g, z, z.copy(), 1)
    assert np.abs(net.get_params() - o.params).max() < 1e-6


def test_training_fits_a_smooth_target():
    """A logistic order-disorder curve in the noise input.  The device training loop reaches the
    loss that the oracle's own Adam loop (same network, same arithmetic) reaches on the same data,
    and the curve is reproduced to a few hundredths."""
    import oracle.mlp as M

    net = make(seed=2)
    rng = np.random.default_rng(3)
    n = 1024
    x = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    t = np.stack([1 / (1 + np.exp(6 * (x[:, 0] - 0.2 * x[:, 1]))), 0.3 * x[:, 2] ** 2, 0.5 * x[:, 3]], axis=1)
    t = t.astype(np.float32)
    p0 = net.get_params()
    l0 = net.loss_grad(x, t)[0]
    epochs, batch = 300, 256
    net.train(x, t, epochs=epochs, batch=batch, seed=4)
    l1 = net.loss_grad(x, t)[0]
    # the oracle's loop: fixed minibatch order, FP32 Adam on the bfloat16-arithmetic gradient
    o = M.MLPOracle(net.dims, p0)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    step = 0
    for ep in range(epochs):
        order = np.random.default_rng(ep).permutation(n)
        for b0 in range(0, n, batch):
            sl = order[b0:b0 + batch]
            _, g = o.loss_grad(x[sl], t[sl])
            step += 1
            m, v = o.adam_step(g, m, v, step)
    lo = o.loss_grad(x, t)[0]
    assert l1 < l0 / 20, (l0, l1)
    assert l1 < 2.0 * lo and lo < 2.0 * l1, (l1, lo)
    y = net.forward(x)
    assert np.sqrt(np.mean((y[:, 0] - t[:, 0]) ** 2)) < 0.05
    assert net.last_train_ms() > 0


def test_binding_checks_shapes():
    """The C side reads as many values as the batch and the layer widths imply, so the binding
    refuses arrays of any other shape; a single point may come as one 1-D row."""
    net = make(n_in=4, hidden=32, n_hidden=2, n_out=3, seed=2)
    x = np.zeros((10, 4), dtype=np.float32)
    t = np.zeros((10, 3), dtype=np.float32)
    for bad in (np.zeros((10, 3)), np.zeros(40)):
        with pytest.raises(ValueError):
            net.forward(bad) if bad.ndim == 2 else net.loss_grad(bad, t)
    with pytest.raises(ValueError):
        net.loss_grad(x, t[:9])
    with pytest.raises(ValueError):
        net.train_step(x, t, w=np.ones(4))
    with pytest.raises(ValueError):
        net.set_params(np.zeros(net.n_params - 1))
    assert net.forward(x[0]).shape == (1, 3)

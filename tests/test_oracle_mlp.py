"""Pins of the surrogate network's CPU oracle (oracle/mlp.py).

* bfloat16 rounding equals torch's float32 -> bfloat16 conversion (round to nearest even),
  including ties and NaN.
* The float64 network's gradient equals central finite differences of its loss.
* The first Adam step moves every parameter by lr * g / (|g| + eps) (bias-corrected moments).
"""
import numpy as np
import pytest


def test_bf16_rounding_matches_torch():
    import torch

    import oracle.mlp as M

    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.array([0.0, -0.0, 1.0, np.inf, -np.inf, 3.3895313892515355e38], dtype=np.float32)])
    # exact ties: mantissa low half 0x8000 with even and odd upper halves
    ties = (np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x40490000 | 0x8000], dtype=np.uint32)).view(np.float32)
    x = np.concatenate([x, ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = M.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isnan(M.bf16_round(np.array([np.nan], dtype=np.float32)))[0]


def test_exact_gradient_matches_finite_differences():
    import oracle.mlp as M

    dims = [4, 16, 16, 3]
    rng = np.random.default_rng(1)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = rng.standard_normal(n) * 0.4
    x = rng.standard_normal((9, 4))
    t = rng.standard_normal((9, 3))
    w = np.array([1.0, 0.5, 2.0])
    o = M.MLPOracle(dims, p, exact=True)
    _, g = o.loss_grad(x, t, w)
    h = 1e-6
    idx = rng.choice(n, 60, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = M.MLPOracle(dims, pp, exact=True).loss_grad(x, t, w)[0]
        lm = M.MLPOracle(dims, pm, exact=True).loss_grad(x, t, w)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    import oracle.mlp as M

    dims = [4, 128, 128, 128, 3]
    rng = np.random.default_rng(2)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = (rng.standard_normal(n) * 0.1).astype(np.float32)
    x = rng.standard_normal((64, 4)).astype(np.float32)
    t = rng.standard_normal((64, 3)).astype(np.float32)
    _, g16 = M.MLPOracle(dims, p).loss_grad(x, t)
    _, g64 = M.MLPOracle(dims, p, exact=True).loss_grad(x, t)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_first_adam_step_closed_form():
    import oracle.mlp as M

    dims = [2, 3, 1]
    n = 2 * 3 + 3 + 3 + 1
    p = np.linspace(-1, 1, n).astype(np.float32)
    g = np.linspace(-0.3, 0.5, n).astype(np.float32)
    o = M.MLPOracle(dims, p)
    o.adam_step(g, np.zeros(n, np.float32), np.zeros(n, np.float32), 1, lr=1e-2)
    want = p - 1e-2 * g / (np.abs(g) + 1e-8)
    assert np.allclose(o.params, want, rtol=1e-5, atol=1e-7)

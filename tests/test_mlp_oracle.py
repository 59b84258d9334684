"""Pins of the MLP oracle (oracle/mlp.py): bfloat16 rounding on known
values, and the analytic gradients of the pairwise ranking loss against
central finite differences of the float64 forward pass."""

from __future__ import annotations

import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    f = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5, 0.0, 65504.0],
                 np.float32)
    got = mlp.bf16(f)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1).  1 + 3*2^-9 rounds up.
    # 1 + 2^-7 + 2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: to even (1 + 2^-6).
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 0.0, 65536.0], np.float32)
    assert np.array_equal(got, want)


def test_gelu_grad_matches_difference():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h)
    assert np.allclose(mlp.gelu_grad(x), fd, atol=1e-8)


def test_pair_loss_known_value():
    s = np.array([0.0, 0.0, 2.0])
    loss, ds = mlp.pair_loss(s, [[0, 1], [2, 0]])
    # softplus(0) = ln 2; softplus(-2) = ln(1 + e^-2)
    assert np.isclose(loss, (np.log(2.0) + np.log1p(np.exp(-2.0))) / 2)
    assert np.isclose(ds.sum(), 0.0)


def test_analytic_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    w = rng.normal(scale=0.3, size=mlp.N_PARAMS)
    norm = np.concatenate([rng.normal(size=16) * 0.1, 1.0 + rng.random(16)])
    x = rng.normal(size=(24, 16))
    pairs = np.array([(i, (i * 7 + 3) % 24) for i in range(24) if i != (i * 7 + 3) % 24])
    loss, g = mlp.gradients(w, norm, x, pairs, bf16_ops=False)
    idx = np.concatenate([rng.choice(np.arange(0, 1024), 30), rng.choice(np.arange(1024, 1088), 10),
                          rng.choice(np.arange(1088, 5184), 40), rng.choice(np.arange(5184, 5248), 10),
                          rng.choice(np.arange(5248, 5312), 10), [5312]])
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        lp = mlp.pair_loss(mlp.forward(wp, norm, x, False)[0], pairs)[0]
        lm = mlp.pair_loss(mlp.forward(wm, norm, x, False)[0], pairs)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_bf16_gradients_close_to_exact():
    """The bf16 path is the exact path up to bfloat16 rounding (8-bit mantissa)."""
    rng = np.random.default_rng(4)
    w = rng.normal(scale=0.3, size=mlp.N_PARAMS)
    norm = np.concatenate([np.zeros(16), np.ones(16)])
    x = rng.normal(size=(64, 16))
    pairs = np.array([(i, (i + 5) % 64) for i in range(64)])
    l64, g64 = mlp.gradients(w, norm, x, pairs, bf16_ops=False)
    l16, g16 = mlp.gradients(w, norm, x, pairs, bf16_ops=True)
    assert abs(l64 - l16) < 0.02 * abs(l64) + 1e-3
    assert np.linalg.norm(g64 - g16) < 0.05 * np.linalg.norm(g64)

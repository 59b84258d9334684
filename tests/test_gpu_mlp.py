"""GPU MLP kernels (MFMA bf16) against the CPU oracle (oracle/mlp.py):
scores, gradients of the pairwise ranking loss, and that Adam steps on the
GPU learn a ranking."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as ref
from rasterizer_amd import Mlp

pytestmark = pytest.mark.gpu


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 16)).astype(np.float32) * 2 + 0.5
    norm = np.concatenate([np.full(16, 0.5), np.full(16, 0.5)]).astype(np.float32)
    return rng, x, norm


@pytest.mark.parametrize("n", [1, 31, 32, 200, 5000])
def test_forward_matches_oracle(n):
    rng, x, norm = _data(n, n)
    m = Mlp(seed=7)
    w = rng.normal(scale=0.25, size=Mlp.N_PARAMS).astype(np.float32)
    m.set_params(w, norm)
    got = m.forward(x)
    want, _ = ref.forward(w, norm, x, bf16_ops=True)
    assert np.allclose(got, want, rtol=2e-3, atol=2e-3 * (1 + np.abs(want).max())), np.abs(got - want).max()


@pytest.mark.parametrize("n", [40, 3000])
def test_gradients_match_oracle(n):
    rng, x, norm = _data(100 + n, n)
    m = Mlp(seed=3)
    w = rng.normal(scale=0.25, size=Mlp.N_PARAMS).astype(np.float32)
    m.set_params(w, norm)
    i = rng.integers(0, n, 4 * n)
    j = rng.integers(0, n, 4 * n)
    pairs = np.stack([i, j], 1)[i != j].astype(np.uint32)
    g, loss = m.gradients(x, pairs)
    lw, gw = ref.gradients(w, norm, x, pairs, bf16_ops=True)
    assert abs(loss - lw) <= 1e-4 * (1 + abs(lw))
    floor = 1e-3 * np.abs(gw).max()  # b3's gradient is a sum of +-pairs: exactly zero up to rounding
    for name, a, b in (("W1", 0, 1024), ("b1", 1024, 1088), ("W2", 1088, 5184), ("b2", 5184, 5248),
                       ("W3", 5248, 5312), ("b3", 5312, 5313)):
        scale = max(np.abs(gw[a:b]).max(), floor)
        err = np.abs(g[a:b] - gw[a:b]).max() / scale
        assert err < 2e-2, f"{name}: max error {err:.3e} of the block's largest gradient"


def test_training_learns_a_ranking():
    rng = np.random.default_rng(0)
    n = 4096
    x = rng.normal(size=(n, 16)).astype(np.float32)
    v = rng.normal(size=16)
    y = np.tanh(x @ v / 4) + 0.1 * np.sin(3 * x[:, 0])
    m = Mlp(seed=1)
    m.set_params(norm=np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32))
    losses = []
    for step in range(300):
        i = rng.integers(0, n, 8192)
        j = rng.integers(0, n, 8192)
        keep = y[i] != y[j]
        hi = np.where(y[i] > y[j], i, j)[keep]
        lo = np.where(y[i] > y[j], j, i)[keep]
        losses.append(m.train_step(x, np.stack([hi, lo], 1).astype(np.uint32), lr=3e-3))
    s = m.forward(x)
    i = rng.integers(0, n, 20000)
    j = rng.integers(0, n, 20000)
    agree = np.mean(np.sign(s[i] - s[j]) == np.sign(y[i] - y[j]))
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}, pair agreement {agree:.3f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
    assert agree > 0.9

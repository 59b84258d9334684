"""CPU oracle of the batch-scoring MLP (test infrastructure only).

Restates rz_mlp.hip in numpy: 16 -> 64 -> 64 -> 1 with exact-erf GELU,
inputs normalised by (x - mean) * invStd, and the mean pairwise ranking loss
softplus(s_j - s_i) over pairs (i, j).  With bf16=True the products take
their operands rounded to bfloat16 (round to nearest even) where the kernels
feed MFMA -- normalised inputs, W1, h1, W2, dz2, dz1 -- and accumulate in
float32; with bf16=False everything is float64, which is what the
finite-difference test of the analytic gradients uses.  Parameter layout as
in rz_kernels.h: W1[16][64] at 0, b1 at 1024, W2[64][64] at 1088, b2 at
5184, W3[64] at 5248, b3 at 5312.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

N_PARAMS = 5313
W1, B1, W2, B2, W3, B3 = 0, 1024, 1088, 5184, 5248, 5312


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    out = (rounded & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def unpack(w: np.ndarray):
    w = np.asarray(w)
    return (w[W1:B1].reshape(16, 64), w[B1:W2], w[W2:B2].reshape(64, 64), w[B2:W3], w[W3:B3], w[B3])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def _q(a, on: bool):
    return bf16(a) if on else a


def forward(w, norm, x, bf16_ops: bool = True):
    """Scores and the activations the backward pass needs."""
    dt = np.float32 if bf16_ops else np.float64
    w1, b1, w2, b2, w3, b3 = (np.asarray(t, dt) for t in unpack(np.asarray(w, dt)))
    xn = ((np.asarray(x, dt) - np.asarray(norm[:16], dt)) * np.asarray(norm[16:], dt)).astype(dt)
    xq = _q(xn, bf16_ops)
    z1 = (xq @ _q(w1, bf16_ops)).astype(dt) + b1
    h1 = gelu(z1).astype(dt)
    h1q = _q(h1, bf16_ops)
    z2 = (h1q @ _q(w2, bf16_ops)).astype(dt) + b2
    h2 = gelu(z2).astype(dt)
    s = (h2 @ w3).astype(dt) + b3
    return s, dict(xq=xq, z1=z1, h1q=h1q, z2=z2, h2=h2, w2=w2, w3=w3)


def pair_loss(s, pairs):
    """Mean softplus(s_j - s_i) and its gradient in s."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    d = s[pairs[:, 0]] - s[pairs[:, 1]]
    loss = np.mean(np.maximum(-d, 0) + np.log1p(np.exp(-np.abs(d))))
    sg = 1.0 / (1.0 + np.exp(d))
    ds = np.zeros_like(s)
    np.add.at(ds, pairs[:, 0], -sg / len(pairs))
    np.add.at(ds, pairs[:, 1], sg / len(pairs))
    return loss, ds


def gradients(w, norm, x, pairs, bf16_ops: bool = True):
    """Loss and its analytic gradient in the 5313 parameters."""
    dt = np.float32 if bf16_ops else np.float64
    s, a = forward(w, norm, x, bf16_ops)
    loss, ds = pair_loss(s, pairs)
    ds = ds.astype(dt)
    g = np.zeros(N_PARAMS, dt)
    g[W3:B3] = ds @ a["h2"]
    g[B3] = ds.sum()
    dz2 = (ds[:, None] * a["w3"][None, :] * gelu_grad(a["z2"])).astype(dt)
    g[B2:W3] = dz2.sum(0)
    dz2q = _q(dz2, bf16_ops)
    g[W2:B2] = (a["h1q"].T @ dz2q).reshape(-1)
    dh1 = (dz2q @ _q(a["w2"], bf16_ops).T).astype(dt)
    dz1 = (dh1 * gelu_grad(a["z1"])).astype(dt)
    g[B1:W2] = dz1.sum(0)
    g[W1:B1] = (a["xq"].T @ _q(dz1, bf16_ops)).reshape(-1)
    return loss, g

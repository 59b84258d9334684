"""Occluder::bake on the device (rz_bake_device) against the host bake
(rz_bake) and the oracle's bake (oracle.c or_bake), all three on the host
that runs the GPU (RSQRTPS differs between CPU vendors, so the k-means order
is compared on the same host): packets, bounds and centre bit for bit, over
box meshes, random quads, degenerate quads (zero normals, NaN after
normalization), a NaN vertex and signed zeros at the bounds."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import Occluder, Rasterizer, scenes

pytestmark = pytest.mark.gpu


def _meshes(seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(40):
        nb = int(rng.integers(1, 30))
        quads = np.concatenate([scenes.box_quads(rng.uniform(-50, 50, 3), rng.uniform(0.2, 8, 3)) for _ in range(nb)])
        out.append(quads)
    for k in range(20):  # random quads, some degenerate
        q = rng.normal(size=(int(rng.integers(1, 200)), 4, 4)).astype(np.float32) * rng.uniform(0.1, 100)
        q[..., 3] = 1.0
        deg = rng.random(q.shape[0]) < 0.2
        q[deg] = q[deg][:, :1]  # all four vertices equal: zero normal
        out.append(q)
    q = scenes.box_quads((0, 0, 0), (1, 1, 1)).copy()
    q[2, 1, 0] = np.nan  # a NaN vertex: ignored by the bounds, packs as INT_MIN
    out.append(q)
    q = scenes.box_quads((1, 1, 1), (1, 1, 1)).copy()  # min x is 0 on two faces
    q[..., 0] = np.where(q[..., 0] == 0, np.float32(-0.0), q[..., 0])
    q[3, :, 0] = np.where(q[3, :, 0] == 0, np.float32(0.0), q[3, :, 0])
    out.append(q)
    meshes = []
    for quads in out:
        quads = scenes.pad_to_batches(np.asarray(quads, np.float32))
        meshes.append(quads.reshape(-1, 4))
    return meshes, rng


@pytest.mark.parametrize("seed", [0, 1])
def test_device_bake_matches_host_bake(oracle_mod, seed):
    meshes, rng = _meshes(seed)
    refs = []
    for m in meshes:
        finite = m[np.isfinite(m).all(axis=1)]
        lo = finite.min(axis=0) - rng.uniform(0, 2, 4).astype(np.float32)
        hi = finite.max(axis=0) + rng.uniform(0, 2, 4).astype(np.float32)
        refs.append((lo, hi))
    R = Rasterizer(64, 64)
    got = R.bake_device(meshes, [r[0] for r in refs], [r[1] for r in refs])
    assert len(got) == len(meshes)
    for i, (m, (lo, hi), g) in enumerate(zip(meshes, refs, got)):
        h = Occluder.bake(m, lo, hi)
        assert np.array_equal(g.packets, h.packets), f"mesh {i}: {int((g.packets != h.packets).sum())} packet words differ"
        for f in ("m_boundsMin", "m_boundsMax", "m_center"):
            a, b = getattr(g, f).view(np.uint32), getattr(h, f).view(np.uint32)
            assert np.array_equal(a, b), f"mesh {i}: {f} {getattr(g, f)} vs {getattr(h, f)}"
        po, bmin, bmax, cen = oracle_mod.bake(m, lo, hi)
        assert np.array_equal(g.packets, po), f"mesh {i}: {int((g.packets != po).sum())} packet words differ from the oracle"
        for f, o in (("m_boundsMin", bmin), ("m_boundsMax", bmax), ("m_center", cen)):
            assert np.array_equal(getattr(g, f).view(np.uint32), o.view(np.uint32)), f"mesh {i}: {f} vs the oracle"


def test_device_bake_rejects_partial_batches():
    from rasterizer_amd import RasterizerError

    R = Rasterizer(64, 64)
    with pytest.raises(RasterizerError):
        R.bake_device([np.zeros((16, 4), np.float32)], [np.zeros(4)], [np.ones(4)])

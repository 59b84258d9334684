"""QuadDecomposition (rz_quad_decompose) against the reference's own
QuadDecomposition::decompose, built from its unmodified source into
oracle/_ref/ (QuadDecomposition.cpp:311-445): the same quads, in the same
order, with the same vertex order inside each, on terrains, meshes with holes
and flipped diagonals (odd cycles in the candidate graph, so the augmenting
paths shrink blossoms), non-manifold and degenerate meshes.  Host only."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import QuadDecomposition, RasterizerError, scenes


@pytest.fixture(scope="module")
def ref(oracle_mod):
    fn = oracle_mod.quad_reference()
    if fn is None:
        pytest.skip("the reference's QuadDecomposition build (oracle/_ref) is absent")
    return fn


def _grid(n: int, rng, amp: float, flip: float, drop: float):
    g = np.arange(n, dtype=np.float64) * 2.0
    X, Z = np.meshgrid(g, g, indexing="ij")
    Y = amp * rng.normal(size=X.shape)
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(n * n)], axis=1).astype(np.float32)
    ij = np.arange(n * n).reshape(n, n)
    a, b, c, d = ij[:-1, :-1].ravel(), ij[1:, :-1].ravel(), ij[1:, 1:].ravel(), ij[:-1, 1:].ravel()
    f = rng.random(a.size) < flip
    t1 = np.where(f[:, None], np.stack([a, b, d], 1), np.stack([a, b, c], 1))
    t2 = np.where(f[:, None], np.stack([b, c, d], 1), np.stack([a, c, d], 1))
    tri = np.stack([t1, t2], 1).reshape(-1, 3)
    keep = rng.random(tri.shape[0]) >= drop
    tri = tri[keep]
    tri = tri[rng.permutation(tri.shape[0])] if rng.random() < 0.5 else tri
    return tri.astype(np.uint32), V


def _check(ref, tri, V, what):
    want = ref(tri, V)
    got = QuadDecomposition.decompose(tri, V)
    assert got.size == want.size, f"{what}: {got.size // 4} quads, reference {want.size // 4}"
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{what}: first difference in quad {bad[0] // 4}"
    return want


@pytest.mark.parametrize("seed", range(40))
def test_random_meshes(ref, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 40))
    tri, V = _grid(n, rng, amp=float(rng.choice([0.0, 0.05, 0.3, 1.0])), flip=float(rng.random()),
                   drop=float(rng.choice([0.0, 0.05, 0.2, 0.4])))
    _check(ref, tri, V, f"seed {seed}")


def test_terrain_and_buildings(ref):
    """The city scene's heightfield (fBm, amplitude 20, 4 octaves) at 128^2
    and a triangulated box set: the configurations' own meshes."""
    hf = scenes.Heightfield(128, 2.0, 20.0, seed=7)
    tri, V = scenes.heightfield_triangles(hf)
    q = _check(ref, tri, V, "terrain")
    assert q.size // 4 >= tri.shape[0] // 2 * 0.9
    tri, V = scenes.triangulated_boxes(np.random.default_rng(3), 50)
    q = _check(ref, tri, V, "boxes")
    assert q.size // 4 == tri.shape[0] // 2  # every face's two triangles pair


def test_icosphere(ref):
    v, f = scenes.icosphere()
    V = np.concatenate([v, np.ones((v.shape[0], 1))], axis=1).astype(np.float32) * np.float32(3.0)
    V[:, 3] = 1.0
    _check(ref, f.astype(np.uint32), V, "icosphere")


def test_degenerate_and_non_manifold(ref):
    rng = np.random.default_rng(11)
    tri, V = _grid(12, rng, amp=0.1, flip=0.3, drop=0.0)
    extra = [
        tri[:40],                                     # every triangle twice: edges shared by four
        np.array([[0, 1, 0], [5, 5, 5], [2, 3, 2]]),  # repeated indices (a triangle holding an edge and its reverse)
        tri[::-1][:30, ::-1],                         # reversed winding: edges in both directions
    ]
    t = np.concatenate([tri] + extra).astype(np.uint32)
    _check(ref, t, V, "duplicates")
    Vn = V.copy()
    Vn[7, 1] = np.nan  # a NaN vertex: its normals are NaN and the merge test passes
    Vn[20, 0] = np.inf
    _check(ref, tri, Vn, "nan/inf")
    flat = V.copy()
    flat[:, 1] = 0.0
    flat[30] = flat[31]  # a zero-area triangle
    _check(ref, tri, flat, "zero area")


def test_edge_inputs(ref):
    V = np.zeros((4, 4), np.float32)
    assert QuadDecomposition.decompose(np.zeros(0, np.uint32), V).size == 0
    one = np.array([0, 1, 2, 3], np.uint32)  # a trailing partial triangle is ignored
    assert np.array_equal(QuadDecomposition.decompose(one, V), ref(one, V))
    with pytest.raises(RasterizerError):
        QuadDecomposition.decompose(np.array([0, 1, 9], np.uint32), V)

"""QuadDecomposition on the device (rz_quad_decompose_device) against the
reference's own QuadDecomposition::decompose, built from its unmodified
source into oracle/_ref/ (QuadDecomposition.cpp:311-445), and against the
host decomposition (rz_quad_decompose): the same quads, in the same order,
with the same vertex order inside each.  The cases are the host tests'
(random grids with flipped diagonals, holes and shuffled triangle order, so
the augmenting paths shrink blossoms; the city terrain; triangulated boxes;
the icosphere; duplicated and reversed triangles, repeated indices, NaN and
infinite vertices, zero-area triangles) and the 512^2 city terrain."""

from __future__ import annotations

import time

import numpy as np
import pytest

from rasterizer_amd import QuadDecomposition, Rasterizer, RasterizerError, scenes
from test_quad import _grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref(oracle_mod):
    fn = oracle_mod.quad_reference()
    if fn is None:
        pytest.skip("the reference's QuadDecomposition build (oracle/_ref) is absent")
    return fn


@pytest.fixture(scope="module")
def rast():
    return Rasterizer(64, 64)


def _check(ref, rast, tri, V, what):
    want = ref(tri, V)
    got = rast.quad_decompose_device(tri, V)
    assert got.size == want.size, f"{what}: {got.size // 4} quads, reference {want.size // 4}"
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{what}: first difference in quad {bad[0] // 4}"
    host = QuadDecomposition.decompose(tri, V)
    assert np.array_equal(host, got), f"{what}: host and device differ"
    return want


@pytest.mark.parametrize("seed", range(40))
def test_random_meshes(ref, rast, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 40))
    tri, V = _grid(n, rng, amp=float(rng.choice([0.0, 0.05, 0.3, 1.0])), flip=float(rng.random()),
                   drop=float(rng.choice([0.0, 0.05, 0.2, 0.4])))
    _check(ref, rast, tri, V, f"seed {seed}")


def test_terrain_boxes_icosphere(ref, rast):
    hf = scenes.Heightfield(128, 2.0, 20.0, seed=7)
    tri, V = scenes.heightfield_triangles(hf)
    _check(ref, rast, tri, V, "terrain")
    tri, V = scenes.triangulated_boxes(np.random.default_rng(3), 50)
    q = _check(ref, rast, tri, V, "boxes")
    assert q.size // 4 == tri.shape[0] // 2
    v, f = scenes.icosphere()
    V = np.concatenate([v, np.ones((v.shape[0], 1))], axis=1).astype(np.float32) * np.float32(3.0)
    V[:, 3] = 1.0
    _check(ref, rast, f.astype(np.uint32), V, "icosphere")


def test_degenerate_and_non_manifold(ref, rast):
    rng = np.random.default_rng(11)
    tri, V = _grid(12, rng, amp=0.1, flip=0.3, drop=0.0)
    extra = [tri[:40], np.array([[0, 1, 0], [5, 5, 5], [2, 3, 2]]), tri[::-1][:30, ::-1]]
    t = np.concatenate([tri] + extra).astype(np.uint32)
    _check(ref, rast, t, V, "duplicates")
    Vn = V.copy()
    Vn[7, 1] = np.nan
    Vn[20, 0] = np.inf
    _check(ref, rast, tri, Vn, "nan/inf")
    flat = V.copy()
    flat[:, 1] = 0.0
    flat[30] = flat[31]
    _check(ref, rast, tri, flat, "zero area")


def test_edge_inputs(ref, rast):
    V = np.zeros((4, 4), np.float32)
    assert rast.quad_decompose_device(np.zeros(0, np.uint32), V).size == 0
    one = np.array([0, 1, 2, 3], np.uint32)  # a trailing partial index is ignored
    assert np.array_equal(rast.quad_decompose_device(one, V), ref(one, V))
    with pytest.raises(RasterizerError):
        rast.quad_decompose_device(np.array([0, 1, 9], np.uint32), V)


def test_city_terrain_512(ref, rast):
    """The configurations' 512^2 city terrain (522,242 triangles): identical
    to the reference build; device and host times reported."""
    hf = scenes.Heightfield(512, 2.0, 20.0, seed=1)
    tri, V = scenes.heightfield_triangles(hf)
    t0 = time.perf_counter()
    want = ref(tri, V)
    t_ref = time.perf_counter() - t0
    rast.quad_decompose_device(tri, V)  # warm-up (first rocPRIM launches)
    t0 = time.perf_counter()
    got = rast.quad_decompose_device(tri, V)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    host = QuadDecomposition.decompose(tri, V)
    t_host = time.perf_counter() - t0
    assert np.array_equal(got, want)
    assert np.array_equal(host, want)
    print(f"QuadDecomposition 512^2 terrain ({tri.shape[0]:,} triangles -> {want.size // 4:,} quads): "
          f"device {t_dev * 1e3:.1f} ms, host {t_host * 1e3:.1f} ms, reference build {t_ref * 1e3:.1f} ms", flush=True)


@pytest.mark.parametrize("which", ["terrain", "boxes"])
def test_mesh_prepared_on_device(rast, which):
    """rz_prepare_mesh_device (Main.cpp:86-128 from the triangle mesh: the
    decomposition, padding, SAH batches, the vertex buffer's box and the bake,
    all in HBM) against the host chain (rz_quad_decompose, then
    scenes.bake_scene with the vertex buffer's box), bit for bit.  An extra
    vertex that no triangle references widens the scene box, as Main.cpp:
    109-113 takes the box over the whole vertex buffer."""
    from test_gpu_sah import _same_occluders

    if which == "terrain":
        tri, V = scenes.heightfield_triangles(scenes.Heightfield(97, 4.0, 20.0, seed=3))
    else:
        tri, V = scenes.triangulated_boxes(np.random.default_rng(5), 300)
    V = np.concatenate([V, np.array([[500.0, -300.0, 250.0, 1.0]], np.float32)])
    want = scenes.bake_scene(scenes.decompose_mesh(tri, V), vertex_buffer=V).occluders
    got = rast.prepare_mesh_device(tri, V)
    _same_occluders(got, want)
    assert float(want[0].m_refMax[0]) == 500.0 and float(want[0].m_refMin[1]) == -300.0


def test_prepare_mesh_arguments(rast):
    tri, V = scenes.triangulated_boxes(np.random.default_rng(6), 4)
    assert rast.prepare_mesh_device(tri[:0], V) == []
    with pytest.raises(RasterizerError):
        rast.prepare_mesh_device(tri, V, 512, 4)  # granularity not a multiple of 8
    bad = np.concatenate([V, np.array([[np.nan, 0.0, 0.0, 1.0]], np.float32)])  # not referenced, still in the box
    with pytest.raises(RasterizerError):
        rast.prepare_mesh_device(tri, bad)
    with pytest.raises(RasterizerError):
        rast.prepare_mesh_device(np.array([[0, 1, V.shape[0]]], np.uint32), V)

"""SurfaceAreaHeuristic::generateBatches on the device (rz_sah_batches)
against scenes.sah_batches (its restatement, pinned by tests/test_sah.py):
the same batches, in the same order, with the same quad order inside each,
over random boxes with many equal centres and signed zeros, a symmetric grid
(ties between axes and between split positions), a scene's quads, sets below
the target and at the granularity limit, and the argument checks."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, RasterizerError, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rast():
    return Rasterizer(64, 64)


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))


def _random(seed, n, grid):
    rng = np.random.default_rng(seed)
    c = rng.integers(-grid, grid + 1, size=(n, 3)).astype(np.float32)
    h = rng.integers(0, 3, size=(n, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    z = rng.random((n, 3)) < 0.1
    mn[z & (mn == 0)] = -0.0
    mx[z & (mx == 0)] = -0.0
    return mn, mx


@pytest.mark.parametrize("seed,n,target,gran,grid", [
    (1, 5000, 512, 8, 20), (2, 5000, 64, 8, 3), (3, 20000, 512, 8, 1000), (4, 3000, 40, 4, 2),
    (5, 1500, 33, 16, 6)])
def test_random_boxes(rast, seed, n, target, gran, grid):
    mn, mx = _random(seed, n, grid)
    print(f"sah random seed={seed} n={n} target={target} gran={gran}", flush=True)
    assert _same(rast.sah_batches_device(mn, mx, target, gran), scenes.sah_batches(mn, mx, target, gran))


@pytest.mark.parametrize("side,target,gran", [(8, 20, 4), (32, 100, 8), (64, 512, 8)])
def test_symmetric_grid(rast, side, target, gran):
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    c = np.stack([i.ravel(), np.zeros(side * side), j.ravel()], axis=1).astype(np.float32) * 3
    mn, mx = c - 1, c + 1
    print(f"sah grid side={side}", flush=True)
    assert _same(rast.sah_batches_device(mn, mx, target, gran), scenes.sah_batches(mn, mx, target, gran))


def test_scene_quads(rast):
    """The quads of a padded box scene, as Main.cpp:96-107 batches them."""
    rng = scenes.SplitMix64(11)
    c, h = scenes.random_boxes(rng, 2048, -50.0, 50.0, 0.5, 5.0)
    quads = scenes.pad_to_batches(np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(2048)], axis=0))
    mn, mx = quads.min(axis=1), quads.max(axis=1)
    print(f"sah scene quads={quads.shape[0]}", flush=True)
    got = rast.sah_batches_device(mn, mx)
    assert _same(got, scenes.sah_batches(mn, mx))
    assert sorted(np.concatenate(got).tolist()) == list(range(quads.shape[0]))


@pytest.mark.parametrize("n", [0, 1, 16, 17, 24, 300, 511, 512, 513])
def test_small_sets(rast, n):
    mn, mx = _random(100 + n, n, 4)
    assert _same(rast.sah_batches_device(mn, mx, 512, 8), scenes.sah_batches(mn, mx, 512, 8))


def test_arguments(rast):
    mn, mx = _random(7, 100, 4)
    for target, gran in ((16, 8), (512, 0)):
        with pytest.raises(RasterizerError):
            rast.sah_batches_device(mn, mx, target, gran)
    for bad in (np.nan, np.inf, 2.0 ** 41):
        m = mn.copy()
        m[37, 1] = bad
        with pytest.raises(RasterizerError):
            rast.sah_batches_device(m, mx, 512, 8)
    m = mn.copy()
    m[5, 2] = mx[5, 2] + 1  # min > max
    with pytest.raises(RasterizerError):
        rast.sah_batches_device(m, mx, 512, 8)


@pytest.mark.parametrize("seed,n,target,gran,grid", [(21, 300, 512, 8, 5), (22, 20000, 512, 8, 1000),
                                                      (23, 5000, 64, 8, 3)])
def test_matches_reference_build(rast, seed, n, target, gran, grid):
    """The device batches against the reference's own generateBatches (built
    from its unmodified source into oracle/_ref/ where the sources were)."""
    import oracle

    ref = oracle.sah_reference()
    if ref is None:
        pytest.skip("oracle/_ref/libsah_ref.so is not built")
    mn, mx = _random(seed, n, grid)
    assert _same(rast.sah_batches_device(mn, mx, target, gran), ref(mn, mx, target, gran))


def _scene_quads(which):
    if which == "boxes":  # 1,501 boxes: 9,006 quads, padded to 9,008
        rng = scenes.SplitMix64(13)
        c, h = scenes.random_boxes(rng, 1501, -50.0, 50.0, 0.5, 5.0)
        return np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(1501)], axis=0)
    return scenes.Heightfield(97, 4.0, 20.0, seed=3).quads()


def _same_occluders(got, want):
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert np.array_equal(a.packets, b.packets)
        assert np.array_equal(a.m_boundsMin, b.m_boundsMin) and np.array_equal(a.m_boundsMax, b.m_boundsMax)
        assert np.array_equal(a.m_center, b.m_center)
        assert np.array_equal(a.m_refMin, b.m_refMin) and np.array_equal(a.m_refMax, b.m_refMax)


@pytest.mark.parametrize("which", ["boxes", "city"])
def test_scene_prepared_entirely_on_device(rast, which):
    """rz_prepare_scene_device (padding, quad boxes, scene box, SAH batches,
    gather and bake in HBM) against the host preparation, bit for bit."""
    quads = _scene_quads(which)
    print(f"prepare_scene_device {which} quads={quads.shape[0]}", flush=True)
    _same_occluders(rast.prepare_scene_device(quads), scenes.bake_scene(quads).occluders)


def test_prepare_scene_arguments(rast):
    quads = _scene_quads("city")[:100]
    for target, gran in ((512, 4), (512, 12), (16, 8)):
        with pytest.raises(RasterizerError):
            rast.prepare_scene_device(quads, target, gran)
    assert rast.prepare_scene_device(quads[:0]) == []
    # Vertices the SAH costs cannot take: NaN, infinite or beyond 2^40.
    for bad in (np.nan, np.inf, -np.inf, 2.0 ** 41):
        q = quads[:40].copy()
        q[17, 2, 1] = bad
        with pytest.raises(RasterizerError) as e:
            rast.prepare_scene_device(q)
        assert e.value.status == -1, bad
    # 5 quads pad to 8, one batch in input order on both sides.
    _same_occluders(rast.prepare_scene_device(quads[:5]), scenes.bake_scene(quads[:5]).occluders)


@pytest.mark.parametrize("which", ["boxes", "city"])
def test_scene_prepared_on_device(rast, which):
    """Main.cpp:86-128 with SAH batching and the bake both on the device
    (Rasterizer.bake_scene_device) against the host preparation
    (scenes.bake_scene): the same occluders, packets, bounds and centres bit
    for bit."""
    if which == "boxes":
        rng = scenes.SplitMix64(13)
        c, h = scenes.random_boxes(rng, 1500, -50.0, 50.0, 0.5, 5.0)
        quads = np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(1500)], axis=0)
    else:
        quads = scenes.Heightfield(97, 4.0, 20.0, seed=3).quads()
    print(f"scene prep {which} quads={quads.shape[0]}", flush=True)
    want = scenes.bake_scene(quads).occluders
    got = rast.bake_scene_device(quads)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert np.array_equal(a.packets, b.packets)
        assert np.array_equal(a.m_boundsMin, b.m_boundsMin) and np.array_equal(a.m_boundsMax, b.m_boundsMax)
        assert np.array_equal(a.m_center, b.m_center)
        assert np.array_equal(a.m_refMin, b.m_refMin) and np.array_equal(a.m_refMax, b.m_refMax)


def test_draw_from_device_prepared_occluders():
    """Occluders prepared on the device (rz_prepare_scene_device) and wrapped
    by rz_occluder_create_device without leaving HBM draw the same
    query-gated frame, bit for bit, as the host-prepared ones."""
    import ctypes
    import math

    from rasterizer_amd import DeviceBuffer, lib

    W, H = 320, 192
    R = Rasterizer(W, H)
    rng = scenes.SplitMix64(8)
    c, h = scenes.random_boxes(rng, 600, -50.0, 50.0, 0.5, 5.0)
    quads = np.ascontiguousarray(np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(600)], axis=0),
                                 dtype=np.float32)
    eye = np.zeros(3, np.float32)
    mvp = scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), math.pi / 3, W / H)
    host = scenes.bake_scene(quads).occluders
    order = scenes.front_to_back(host, eye)

    n = quads.shape[0]
    n_pad = (n + 7) // 8 * 8
    max_occ = n_pad // 8
    offs = np.zeros(max_occ + 1, np.uint32)
    ref = np.zeros(8, np.float32)
    cnt = ctypes.c_uint32(0)
    d_v = DeviceBuffer(R, quads.nbytes, quads)
    d_p = DeviceBuffer(R, n_pad * 16)
    d_b = DeviceBuffer(R, max_occ * 48)
    handles = []
    try:
        assert lib().rz_prepare_scene_device(R.ctx, n, d_v.ptr, 512, 8, d_p.ptr, d_b.ptr, offs.ctypes.data,
                                             ctypes.byref(cnt), ref.ctypes.data) == 0
        k = cnt.value
        assert k == len(host)
        bounds = d_b.read(np.float32, max_occ * 12).reshape(max_occ, 12)
        for i in range(k):
            bmin, bmax = bounds[i, 0:4].copy(), bounds[i, 4:8].copy()
            rmin, rmax = ref[0:4].copy(), ref[4:8].copy()
            hnd = ctypes.c_void_p()
            start = ctypes.c_void_p(d_p.ptr.value + 16 * int(offs[i]))
            assert lib().rz_occluder_create_device(R.ctx, start, (int(offs[i + 1]) - int(offs[i])) // 2,
                                                   rmin.ctypes.data, rmax.ctypes.data, bmin.ctypes.data,
                                                   bmax.ctypes.data, ctypes.byref(hnd)) == 0
            handles.append(hnd)
        zeros_d = np.zeros(R.blocks_x * R.blocks_y * 64, np.uint16)
        zeros_h = np.zeros(R.blocks_x * R.blocks_y + 8, np.uint16)
        R.setModelViewProjection(mvp)

        R.upload_raw(zeros_d, zeros_h)
        R.clear()
        arr = (ctypes.c_void_p * k)(*[handles[j].value for j in order])
        R.draw_frame_handles(arr, k)
        R.sync()
        dev_depth, dev_hiz = R.raw()

        R.upload_raw(zeros_d, zeros_h)
        R.clear()
        R.draw_frame([host[j] for j in order])
        host_depth, host_hiz = R.raw()
        assert np.array_equal(dev_depth, host_depth) and np.array_equal(dev_hiz, host_hiz)
        assert np.count_nonzero(host_hiz[:-8] != 1) > 0  # something was drawn
    finally:
        R.sync()
        for hnd in handles:
            lib().rz_occluder_destroy(hnd)
        for b in (d_v, d_p, d_b):
            b.free()


def test_one_million_quads(rast):
    """1,000,000 random quad boxes at (512, 8): a partition into batches below
    the target, and, where the reference's own build is present, its exact
    batches (about 4 s of single-threaded reference work)."""
    import oracle

    n = 1_000_000
    rng = np.random.default_rng(n)
    c = rng.uniform(-1000, 1000, (n, 3)).astype(np.float32)
    h = np.exp(rng.uniform(np.log(0.5), np.log(5), (n, 3))).astype(np.float32)
    mn, mx = c - h, c + h
    print("sah 1M quads", flush=True)
    got = rast.sah_batches_device(mn, mx)
    sizes = np.array([b.size for b in got])
    assert np.all((sizes > 0) & (sizes < 512) & (sizes % 8 == 0))
    assert np.array_equal(np.sort(np.concatenate(got)), np.arange(n))
    ref = oracle.sah_reference()
    if ref is not None:
        assert _same(got, ref(mn, mx, 512, 8))


@pytest.mark.parametrize("seed,n,target,gran,grid", [(31, 20000, 512, 8, 1000), (32, 300, 512, 8, 5)])
def test_cpp_dropin_header(tmp_path, seed, n, target, gran, grid):
    """SurfaceAreaHeuristic::generateBatches through the C++ drop-in header
    (include/rasterizer_amd/SurfaceAreaHeuristic.h), called from a compiled
    program as Main.cpp:107 calls it, against the restatement and, where it
    is built, the reference's own generateBatches."""
    import subprocess
    from pathlib import Path

    import oracle

    exe = Path(__file__).resolve().parent / "_bin" / "sah_dropin_check"
    if not exe.exists():
        pytest.fail("tests/_bin/sah_dropin_check is not built (python -m rasterizer_amd.build)")
    mn, mx = _random(seed, n, grid)
    boxes = np.zeros((n, 8), np.float32)
    boxes[:, 0:3], boxes[:, 4:7] = mn, mx
    src, dst = tmp_path / "boxes.f32", tmp_path / "batches.u32"
    boxes.tofile(src)
    print(f"sah drop-in n={n}", flush=True)
    r = subprocess.run([str(exe), str(src), str(dst), str(target), str(gran)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    words = np.fromfile(dst, np.uint32)
    got, at = [], 1
    for _ in range(int(words[0])):
        size = int(words[at])
        got.append(words[at + 1:at + 1 + size].astype(np.int64))
        at += 1 + size
    assert at == words.size
    assert _same(got, scenes.sah_batches(mn, mx, target, gran))
    ref = oracle.sah_reference()
    if ref is not None:
        assert _same(got, ref(mn, mx, target, gran))


def test_python_class_mirror():
    """rasterizer_amd.SurfaceAreaHeuristic.generateBatches, named as the
    reference's class, on Aabb-layout boxes."""
    from rasterizer_amd import SurfaceAreaHeuristic

    mn, mx = _random(41, 3000, 6)
    boxes = np.zeros((3000, 2, 4), np.float32)
    boxes[:, 0, :3], boxes[:, 1, :3] = mn, mx
    got = SurfaceAreaHeuristic.generateBatches(boxes, 512, 8)
    want = scenes.sah_batches(mn, mx, 512, 8)
    assert [list(map(int, b)) for b in want] == got

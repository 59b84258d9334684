"""SAH batching restatement (scenes.sah_batches, SurfaceAreaHeuristic.cpp:10-104)
pinned against an independent brute force, a known answer and its invariants."""
import numpy as np
import pytest

from rasterizer_amd import scenes

f32 = np.float32


def _area(lo, hi):
    # Aabb::surfaceArea (VectorMath.h:57-62) one scalar operation at a time.
    e = [f32(hi[j] - lo[j]) for j in range(3)]
    return f32(f32(f32(e[0] * e[1]) + f32(e[1] * e[2])) + f32(f32(e[2] * e[0]) + f32(0.0)))


def _box(mins, maxs, ids):
    return ([min(mins[i][j] for i in ids) for j in range(3)], [max(maxs[i][j] for i in ids) for j in range(3)])


def _brute(mins, maxs, target, gran):
    """generateBatches with Python's stable sorted() and, for every candidate
    split, both boxes built from scratch over their items (no prefix scans)."""
    centre = [[f32(mins[i][j] + maxs[i][j]) for j in range(3)] for i in range(len(mins))]
    out = []

    def split(ids):
        n = len(ids)
        best, best_axis, best_k = f32(np.inf), -1, -1
        for axis in range(3):
            ids = sorted(ids, key=lambda i: centre[i][axis])
            for k in range(gran, n - gran, gran):
                cost = f32(f32(_area(*_box(mins, maxs, ids[:k])) * f32(k)) +
                           f32(_area(*_box(mins, maxs, ids[k:])) * f32(n - k)))
                if cost < best:
                    best, best_axis, best_k = cost, axis, k
        ids = sorted(ids, key=lambda i: centre[i][best_axis])
        for part in (ids[:best_k], ids[best_k:]):
            if len(part) < target:
                out.append(part)
            else:
                split(part)

    split(list(range(len(mins))))
    return out


def _boxes(seed, n, grid):
    rng = np.random.default_rng(seed)
    # Coarse integer centres make equal keys common; some zeros carry a sign.
    c = rng.integers(-grid, grid + 1, size=(n, 3)).astype(np.float32)
    h = rng.integers(0, 3, size=(n, 3)).astype(np.float32)
    mn, mx = c - h, c + h
    z = rng.random((n, 3)) < 0.1
    mn[z & (mn == 0)] = -0.0
    mx[z & (mx == 0)] = -0.0
    return mn, mx


@pytest.mark.parametrize("seed,n,target,gran,grid", [
    (1, 40, 12, 4, 3), (2, 96, 20, 4, 2), (3, 150, 33, 8, 5), (4, 200, 40, 6, 50), (5, 64, 30, 8, 1),
    (6, 23, 12, 4, 2)])
def test_restatement_matches_brute_force(seed, n, target, gran, grid):
    mn, mx = _boxes(seed, n, grid)
    got = [list(map(int, b)) for b in scenes.sah_batches(mn, mx, target, gran)]
    want = _brute(mn.tolist(), mx.tolist(), target, gran)
    want = [[int(i) for i in b] for b in want]
    assert got == want


@pytest.mark.parametrize("side,target,gran", [(8, 20, 4), (6, 12, 2), (10, 40, 8)])
def test_restatement_ties_on_symmetric_grid(side, target, gran):
    # Unit boxes on a side x side grid in the xz plane, listed row by row:
    # swapping x and z maps the set onto itself, so x and z splits cost
    # exactly the same, and mirror-image split positions tie as well.  Only
    # the first strict minimum over (axis, position) gives the reference's
    # batches.
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    c = np.stack([i.ravel(), np.zeros(side * side), j.ravel()], axis=1).astype(np.float32) * 3
    mn, mx = c - 1, c + 1
    got = [list(map(int, b)) for b in scenes.sah_batches(mn, mx, target, gran)]
    assert got == _brute(mn.tolist(), mx.tolist(), target, gran)


def test_two_clusters_known_answer():
    # 64 unit boxes near x = 0 and 64 near x = 1000, interleaved in the input:
    # the only cheap split separates the clusters, the near one first, and
    # both halves are below the target.
    n = 128
    rng = np.random.default_rng(9)
    c = rng.uniform(0, 10, size=(n, 3)).astype(np.float32)
    far = np.arange(n) % 2 == 1
    c[far, 0] += 1000
    mn, mx = c - 0.5, c + 0.5
    batches = scenes.sah_batches(mn, mx, 100, 8)
    assert len(batches) == 2
    assert sorted(batches[0].tolist()) == np.flatnonzero(~far).tolist()
    assert sorted(batches[1].tolist()) == np.flatnonzero(far).tolist()
    # Each batch is in x order (the final stable sort along the winning axis).
    for b in batches:
        x = (mn[b, 0] + mx[b, 0]).astype(np.float32)
        assert np.all(np.diff(x) >= 0)


def test_invariants():
    rng = np.random.default_rng(3)
    c = rng.uniform(-50, 50, size=(4096, 3)).astype(np.float32)
    h = np.exp(rng.uniform(np.log(0.5), np.log(5), size=(4096, 3))).astype(np.float32)
    mn, mx = (c - h).astype(np.float32), (c + h).astype(np.float32)
    batches = scenes.sah_batches(mn, mx, 512, 8)
    flat = np.concatenate(batches)
    assert sorted(flat.tolist()) == list(range(4096))
    assert all(0 < b.size < 512 and b.size % 8 == 0 for b in batches)


def test_small_sets_and_arguments():
    mn = np.zeros((20, 3), np.float32)
    mx = np.ones((20, 3), np.float32)
    assert scenes.sah_batches(mn[:0], mx[:0], 512, 8) == []
    one = scenes.sah_batches(mn[:16], mx[:16], 512, 8)
    assert len(one) == 1 and one[0].tolist() == list(range(16))
    # Below the target the whole set is still split once, as generateBatches does.
    two = scenes.sah_batches(mn, mx, 512, 8)
    assert [b.size for b in two] == [8, 12]
    with pytest.raises(ValueError):
        scenes.sah_batches(mn, mx, 16, 8)
    with pytest.raises(ValueError):
        scenes.sah_batches(mn, mx, 16, 0)


def _reference_or_skip():
    import oracle

    ref = oracle.sah_reference()
    if ref is None:
        pytest.skip("oracle/_ref/libsah_ref.so is not built (the reference's sources are not on this machine)")
    return ref


@pytest.mark.parametrize("seed,n,target,gran,grid", [
    (11, 40, 12, 4, 3), (12, 300, 512, 8, 5), (13, 5000, 512, 8, 20), (14, 5000, 64, 8, 3),
    (15, 20000, 512, 8, 1000), (16, 3000, 40, 4, 2), (17, 1500, 33, 16, 6), (18, 17, 512, 8, 2)])
def test_restatement_matches_reference_build(seed, n, target, gran, grid):
    """scenes.sah_batches against the reference's own generateBatches, built
    from its unmodified source by oracle/Makefile."""
    ref = _reference_or_skip()
    mn, mx = _boxes(seed, n, grid)
    want = ref(mn, mx, target, gran)
    got = scenes.sah_batches(mn, mx, target, gran)
    assert len(got) == len(want) and all(np.array_equal(a, b) for a, b in zip(got, want))


@pytest.mark.parametrize("side,target,gran", [(8, 20, 4), (32, 100, 8), (64, 512, 8)])
def test_symmetric_grid_matches_reference_build(side, target, gran):
    ref = _reference_or_skip()
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    c = np.stack([i.ravel(), np.zeros(side * side), j.ravel()], axis=1).astype(np.float32) * 3
    mn, mx = c - 1, c + 1
    want = ref(mn, mx, target, gran)
    got = scenes.sah_batches(mn, mx, target, gran)
    assert len(got) == len(want) and all(np.array_equal(a, b) for a, b in zip(got, want))


def test_scene_quads_match_reference_build():
    """A padded box scene and a terrain heightfield, batched as Main.cpp:96-107
    does (target 512, granularity 8)."""
    ref = _reference_or_skip()
    rng = scenes.SplitMix64(11)
    c, h = scenes.random_boxes(rng, 2048, -50.0, 50.0, 0.5, 5.0)
    boxes = scenes.pad_to_batches(np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(2048)], axis=0))
    terrain = scenes.pad_to_batches(scenes.Heightfield(129, 4.0, 20.0, seed=5).quads())
    for quads in (boxes, terrain):
        mn, mx = quads.min(axis=1), quads.max(axis=1)
        want = ref(mn, mx, 512, 8)
        got = scenes.sah_batches(mn, mx, 512, 8)
        assert len(got) == len(want) and all(np.array_equal(a, b) for a, b in zip(got, want))

"""GPU parity of the device-side front-to-back order (rz_views_set_eye):
each view's occluders, submitted in any order, are drawn sorted by the
squared distance from the view's eye to their centre (Main.cpp:184-190,
ties in submission order), and the view equals the oracle drawing them in
that order, with statuses and query bits reported in submission order."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, RasterizerError, Views, scenes

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (320, 192), (256, 256), (1280, 720), (200, 104), (1920, 1080)]


@pytest.fixture(scope="module")
def scene():
    return scenes.config1(seed=4, n_boxes=4096)


def _setup(scene, step: int):
    """Per view: eye, MVP, occluder subset in a shuffled submission order,
    queries, and whether the view is sorted on the device."""
    rng = np.random.default_rng(300 + step)
    out = []
    for k, (w, h) in enumerate(SIZES):
        eye = rng.uniform(-60, 60, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        up = (0, 1, 0) if abs(d[1]) < 0.95 else (1, 0, 0)
        mvp = scenes.view_projection(eye, d, up, np.pi / 3, w / h)
        sel = np.nonzero(rng.random(len(scene.occluders)) < rng.uniform(0.3, 1.0))[0]
        occ = [scene.occluders[i] for i in rng.permutation(sel)]
        q = scene.meta["queries"][rng.choice(len(scene.meta["queries"]), 200 + 50 * k, replace=False)]
        out.append((eye, mvp, occ, q, k % 3 != 2))
    return out


def _run_and_check(oracle_mod, V, setup, gated, report, what, oracles):
    for k, (eye, mvp, occ, q, sorted_) in enumerate(setup):
        (w, h) = SIZES[k]
        order = scenes.front_to_back(occ, eye) if sorted_ else np.arange(len(occ))
        got = V.order(k)
        assert np.array_equal(got, order), f"{what} view {k}: device order differs from front_to_back"
        o = oracles[k]
        o.clear()
        o.setModelViewProjection(mvp)
        st_sorted = o.draw_frame([occ[i] for i in order], gated=gated)
        st_o = np.empty_like(st_sorted)
        st_o[order] = st_sorted
        st_g = V.status(k)
        assert np.array_equal(st_g, st_o), f"{what} view {k}: statuses differ at {np.nonzero(st_g != st_o)[0][:10]}"
        dg, hg = V.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hg, ho), f"{what} view {k}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg, do), f"{what} view {k}: depth {report(dg, do)}"
        vg, cg = V.queries(k)
        vo, co = o.query_batch(q)
        assert np.array_equal(vg, vo) and np.array_equal(cg, co), f"{what} view {k}: query bits differ"


@pytest.mark.parametrize("gated", [False, True])
def test_front_to_back_on_device_matches_oracle(oracle_mod, scene, report, gated):
    """Two steps with new eyes and matrices (graph replays in the classify
    mode), two of every three views sorted on the device."""
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    V.set_query_gated(gated)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in SIZES]
    for step in range(2):
        setup = _setup(scene, step)
        for k, (eye, mvp, occ, q, sorted_) in enumerate(setup):
            V.set_mvp(k, mvp)
            V.set_occluders(k, occ)
            V.set_queries(k, q)
            V.set_eye(k, eye if sorted_ else None)
        V.run(graph=not gated)
        V.sync()
        _run_and_check(oracle_mod, V, setup, gated, report, f"gated={gated} step {step}", oracles)
    # A new eye alone (same plan, same occluder lists): the graph replays with it.
    rng = np.random.default_rng(9)
    setup = [(rng.uniform(-60, 60, 3).astype(np.float32), mvp, occ, q, s) for (_, mvp, occ, q, s) in setup]
    for k, (eye, _, _, _, sorted_) in enumerate(setup):
        if sorted_:
            V.set_eye(k, eye)
    V.run(graph=not gated)
    V.sync()
    _run_and_check(oracle_mod, V, setup, gated, report, f"gated={gated} new eyes", oracles)


def test_front_to_back_ties_keep_submission_order(oracle_mod, scene, report):
    """Occluders at equal distance (the same occluder submitted several
    times) keep their submission order."""
    owner = Rasterizer(64, 64)
    V = Views(owner, [(256, 256)])
    eye = np.array([5.0, 3.0, -40.0], np.float32)
    mvp = scenes.view_projection(eye, np.array([0, 0, 1.0]), (0, 1, 0), np.pi / 3, 1.0)
    m = min(40, len(scene.occluders))
    base = scene.occluders[:m]
    occ = [base[i % m] for i in range(3 * m)]
    V.set_mvp(0, mvp)
    V.set_occluders(0, occ)
    V.set_eye(0, eye)
    V.run(graph=False)
    V.sync()
    got = V.order(0)
    want = scenes.front_to_back(occ, eye)
    assert np.array_equal(got, want)
    for i in range(m):  # the three copies of occluder i appear in submission order
        pos = [int(np.nonzero(got == i + m * c)[0][0]) for c in range(3)]
        assert pos == sorted(pos)


def test_front_to_back_rejects_learned_model(scene):
    from rasterizer_amd import Mlp

    owner = Rasterizer(64, 64)
    V = Views(owner, [(64, 64)])
    V.set_occluders(0, scene.occluders[:10])
    V.set_eye(0, np.zeros(3, np.float32))
    V.set_model(Mlp(seed=1))
    with pytest.raises(RasterizerError):
        V.run(graph=False)


def test_front_to_back_distance_sums_as_dpps():
    """The key is (dx^2 + dy^2) + dz^2 in float32: for d = (1, 2^-12, 2^-12)
    that is exactly 1 (1 + 2^-24 rounds to 1 twice), tying with d = (1, 0, 0)
    so submission order decides; summed as dx^2 + (dy^2 + dz^2), or with a
    fused multiply-add, it would exceed 1 and sort after."""
    from rasterizer_amd import Occluder

    owner = Rasterizer(64, 64)
    V = Views(owner, [(64, 64)])
    packets = np.zeros(32, np.uint32)
    t = np.float32(2.0 ** -12)
    a = Occluder(packets, [0, 0, 0, 1], [1, 1, 1, 1], [1, t, t, 1], [1, t, t, 1])
    b = Occluder(packets, [0, 0, 0, 1], [1, 1, 1, 1], [1, 0, 0, 1], [1, 0, 0, 1])
    assert scenes.front_to_back([a, b], np.zeros(3, np.float32)).tolist() == [0, 1]
    V.set_occluders(0, [a, b])
    V.set_eye(0, np.zeros(3, np.float32))
    V.run(graph=False)
    V.sync()
    assert V.order(0).tolist() == [0, 1]
    V.set_occluders(0, [b, a])
    V.run(graph=False)
    V.sync()
    assert V.order(0).tolist() == [0, 1]

"""GPU parity of batched multi-view launches (rz_views_*): every view equals
the oracle's classify-then-draw sequence on one buffer of its size (clear,
setModelViewProjection, per occluder the depth-independent part of
queryVisibility and rasterize<needsClipping>, then the view's queries), bit
for bit, for graph replays and direct launches alike."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, Views, scenes

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (320, 192), (256, 256), (1280, 720), (200, 104), (16, 8), (512, 512), (8, 8), (1920, 1080)]


@pytest.fixture(scope="module")
def scene():
    return scenes.config1(seed=2, n_boxes=768)


def _views_setup(scene, step: int):
    """Per view: eye, MVP, occluder subset in front-to-back order, queries."""
    rng = np.random.default_rng(100 + step)
    out = []
    for k, (w, h) in enumerate(SIZES):
        eye = rng.uniform(-60, 60, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        up = (0, 1, 0) if abs(d[1]) < 0.95 else (1, 0, 0)
        mvp = scenes.view_projection(eye, d, up, np.pi / 3, w / h)
        sel = np.nonzero(rng.random(len(scene.occluders)) < rng.uniform(0.2, 1.0))[0]
        occ = [scene.occluders[i] for i in sel]
        occ = [occ[i] for i in scenes.front_to_back(occ, eye)]
        q = scene.meta["queries"][rng.choice(len(scene.meta["queries"]), 300 + 100 * k, replace=False)]
        out.append((mvp, occ, q))
    return out


def _check_view(oracle_mod, V, view, w, h, mvp, occ, q, report, what, o=None):
    """o: the view's persistent oracle (clear leaves depth stale, so a buffer
    reused across steps is compared with an oracle reused the same way); with
    a fresh oracle only written blocks (hiZ != 1) are compared for depth."""
    fresh = o is None
    if fresh:
        o = oracle_mod.OracleRasterizer(w, h)
    o.clear()
    o.setModelViewProjection(mvp)
    st_o = o.draw_frame(occ, gated=False)
    st_g = V.status(view)
    assert np.array_equal(st_g, st_o), f"{what}: statuses differ at {np.nonzero(st_g != st_o)[0][:10]}"
    dg, hg = V.raw(view)
    do, ho = o.raw()
    assert np.array_equal(hg, ho), f"{what}: hiZ {report(hg, ho)}"
    if fresh:
        written = np.repeat(ho[: (w // 8) * (h // 8)] != 1, 64)
        dg, do = dg[written], do[written]
    assert np.array_equal(dg, do), f"{what}: depth {report(dg, do)}"
    vg, cg = V.queries(view)
    vo, co = o.query_batch(q)
    assert np.array_equal(vg, vo), f"{what}: {int((vg != vo).sum())} query bits differ"
    assert np.array_equal(cg, co), f"{what}: {int((cg != co).sum())} clip bits differ"
    return int((st_o > 0).sum())


def test_views_match_oracle(oracle_mod, scene, report):
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in SIZES]
    drawn = 0
    for step, graph in enumerate([True, False, True]):
        setup = _views_setup(scene, step)
        for k, (mvp, occ, q) in enumerate(setup):
            V.set_mvp(k, mvp)
            V.set_occluders(k, occ)
            V.set_queries(k, q)
        print(f"step {step} graph={graph}: {V.info()}", flush=True)
        V.run(graph=graph)
        V.sync()
        for k, (w, h) in enumerate(SIZES):
            mvp, occ, q = setup[k]
            drawn += _check_view(oracle_mod, V, k, w, h, mvp, occ, q, report, f"step {step} view {k} {w}x{h}",
                                 oracles[k])
    assert drawn > 20


def test_views_graph_replays_match_direct(oracle_mod, scene, report):
    """Replays of one captured graph with new matrices each time, with direct
    launches and host synchronisation between them: every replay equals the
    direct launch of the same step and the oracle."""
    owner = Rasterizer(64, 64)
    V = Views(owner, SIZES[:6])
    setup = _views_setup(scene, 7)
    for k in range(6):
        V.set_occluders(k, setup[k][1])
        V.set_queries(k, setup[k][2])
    rng = np.random.default_rng(5)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in SIZES[:6]]
    for rep in range(4):
        mvps = []
        for k, (w, h) in enumerate(SIZES[:6]):
            eye = rng.uniform(-40, 40, 3).astype(np.float32)
            mvps.append(scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), np.pi / 3, w / h))
            V.set_mvp(k, mvps[k])
        V.run(graph=True)
        V.sync()
        assert V.info()["graph"] == 1
        replay = [V.raw(k) + V.queries(k) for k in range(6)]
        V.run(graph=False)
        V.sync()
        for k in range(6):
            direct = V.raw(k) + V.queries(k)
            for a, b in zip(replay[k], direct):
                assert np.array_equal(a, b), f"replay {rep} view {k} differs from the direct launch"
        for k, (w, h) in enumerate(SIZES[:6]):
            # The oracle mirrors both steps: the replay, then the direct launch.
            o = oracles[k]
            o.clear()
            o.setModelViewProjection(mvps[k])
            o.draw_frame(setup[k][1], gated=False)
            _check_view(oracle_mod, V, k, w, h, mvps[k], setup[k][1], setup[k][2], report, f"replay {rep} view {k}",
                        o)


def test_views_gated_graph_replays(oracle_mod, scene, report):
    """Query-gated steps replayed from their captured graphs (the launches
    before the cooperative raster launch and the queries after it) with new
    matrices each time, with direct launches between them: every replay
    equals the direct launch of the same step and the oracle's gated frame
    loop."""
    sizes = [s for s in SIZES[:6] if (s[0] // 64 + (s[0] % 64 > 0)) * (s[1] // 64 + (s[1] % 64 > 0)) <= 510]
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    V.set_query_gated(True)
    setup = _views_setup(scene, 9)
    for k in range(len(sizes)):
        V.set_occluders(k, setup[k][1])
        V.set_queries(k, setup[k][2])
    rng = np.random.default_rng(6)
    # Clear leaves depth stale, so each view's oracle is reused across steps
    # and mirrors both launches of a repetition.
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in sizes]
    for rep in range(3):
        mvps = []
        for k, (w, h) in enumerate(sizes):
            eye = rng.uniform(-40, 40, 3).astype(np.float32)
            mvps.append(scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), np.pi / 3, w / h))
            V.set_mvp(k, mvps[k])
        V.run(graph=True)
        V.sync()
        assert V.info()["graph"] == 1
        replay = [V.raw(k) + V.queries(k) + (V.status(k),) for k in range(len(sizes))]
        V.run(graph=False)
        V.sync()
        for k, (w, h) in enumerate(sizes):
            direct = V.raw(k) + V.queries(k) + (V.status(k),)
            for a, b in zip(replay[k], direct):
                assert np.array_equal(a, b), f"gated replay {rep} view {k} differs from the direct launch"
            o = oracles[k]
            for _ in range(2):
                o.clear()
                o.setModelViewProjection(mvps[k])
                st_o = o.draw_frame(setup[k][1], gated=True)
            assert np.array_equal(V.status(k), st_o), f"gated replay {rep} view {k}: statuses"
            dg, hg = V.raw(k)
            do, ho = o.raw()
            assert np.array_equal(hg, ho), f"gated replay {rep} view {k}: hiZ {report(hg, ho)}"
            assert np.array_equal(dg, do), f"gated replay {rep} view {k}: depth {report(dg, do)}"
            vg, cg = V.queries(k)
            vo, co = o.query_batch(setup[k][2])
            assert np.array_equal(vg, vo) and np.array_equal(cg, co)


def test_views_empty_lists():
    owner = Rasterizer(64, 64)
    V = Views(owner, [(64, 64), (32, 16)])
    V.set_queries(1, np.array([[-1, -1, 5, 1, 1, 1, 6, 1]], np.float32))
    V.run(graph=True)
    V.sync()
    d, h = V.raw(0)
    assert np.all(h[: 64] == 1) and np.all(d == 0)
    vis, clip = V.queries(1)
    assert vis.shape == (1,)


def test_views_query_gated_match_draw_frame(oracle_mod, scene, report):
    """Query-gated views: each view equals the reference's frame loop
    (clear, setModelViewProjection, queryVisibility-gated rasterize of every
    occluder in order), as the oracle's draw_frame, over two steps."""
    sizes = [s for s in SIZES if (s[0] // 64 + (s[0] % 64 > 0)) * (s[1] // 64 + (s[1] % 64 > 0)) <= 510]
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    V.set_query_gated(True)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in sizes]
    drawn = skipped = 0
    for step in range(2):
        setup = _views_setup(scene, 20 + step)
        for k, (mvp, occ, q) in enumerate(setup[: len(sizes)]):
            V.set_mvp(k, mvp)
            V.set_occluders(k, occ)
            V.set_queries(k, q)
        print(f"gated step {step}: {V.info()}", flush=True)
        V.run(graph=True)
        V.sync()
        for k, (w, h) in enumerate(sizes):
            mvp, occ, q = setup[k]
            o = oracles[k]
            o.clear()
            o.setModelViewProjection(mvp)
            st_o = o.draw_frame(occ, gated=True)
            st_g = V.status(k)
            assert np.array_equal(st_g, st_o), f"step {step} view {k}: statuses differ at {np.nonzero(st_g != st_o)[0][:10]}"
            dg, hg = V.raw(k)
            do, ho = o.raw()
            assert np.array_equal(hg, ho), f"step {step} view {k}: hiZ {report(hg, ho)}"
            assert np.array_equal(dg, do), f"step {step} view {k}: depth {report(dg, do)}"
            vg, cg = V.queries(k)
            vo, co = o.query_batch(q)
            assert np.array_equal(vg, vo) and np.array_equal(cg, co)
            drawn += int((st_o > 0).sum())
            skipped += int((st_o == 0).sum())
    print(f"gated: {drawn} drawn, {skipped} not drawn", flush=True)
    assert drawn > 20 and skipped > 20


def test_views_query_gated_limits(scene):
    from rasterizer_amd import Mlp, RasterizerError

    owner = Rasterizer(64, 64)
    # 8192x4096: even its 16x16-block tiles (64 x 32) cannot all be
    # co-resident: refused, not run.
    V = Views(owner, [(8192, 4096)])
    V.set_query_gated(True)
    V.set_occluders(0, scene.occluders[:10])
    with pytest.raises(RasterizerError):
        V.run(graph=False)
    # A batch-granular model would query per batch: refused.
    V2 = Views(owner, [(320, 192)])
    V2.set_occluders(0, scene.occluders[:10])
    V2.set_query_gated(True)
    V2.set_model(Mlp(seed=1))
    with pytest.raises(RasterizerError):
        V2.run(graph=False)
    V2.set_model(None)
    V2.run(graph=False)
    V2.sync()


def test_views_query_gated_large_views(oracle_mod, scene, report):
    """Query-gated 2048^2 and 3840x2160 views, whose 8x8-block tiles cannot
    all be co-resident: the step moves to 16x16-block tiles and each view
    still equals the reference's frame loop (the oracle's draw_frame), over
    two steps (the second from the buffers the first left)."""
    sizes = [(2048, 2048), (3840, 2160), (1920, 1080)]
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    V.set_query_gated(True)
    oracles = [oracle_mod.OracleRasterizer(w, h) for w, h in sizes]
    rng = np.random.default_rng(31)
    drawn = skipped = 0
    for step in range(2):
        want = []
        for k, (w, h) in enumerate(sizes):
            eye = rng.uniform(-40, 40, 3).astype(np.float32)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            up = (0, 1, 0) if abs(d[1]) < 0.95 else (1, 0, 0)
            mvp = scenes.view_projection(eye, d, up, np.pi / 3, w / h)
            occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
            V.set_mvp(k, mvp)
            V.set_occluders(k, occ)
            oracles[k].clear()
            oracles[k].setModelViewProjection(mvp)
            st_o = oracles[k].draw_frame(occ, gated=True)
            want.append(st_o)
            drawn += int((st_o > 0).sum())
            skipped += int((st_o == 0).sum())
        print(f"gated large views step {step}: {V.info()}", flush=True)
        V.run(graph=False)
        V.sync()
        for k, (w, h) in enumerate(sizes):
            assert np.array_equal(V.status(k), want[k]), f"step {step} view {k} {w}x{h}: statuses"
            dg, hg = V.raw(k)
            do, ho = oracles[k].raw()
            assert np.array_equal(hg, ho), f"step {step} view {k} {w}x{h}: hiZ {report(hg, ho)}"
            assert np.array_equal(dg, do), f"step {step} view {k} {w}x{h}: depth {report(dg, do)}"
    assert drawn > 20 and skipped > 20


def test_stress_scene_views_match_oracle(oracle_mod, report):
    """The stress configuration at reduced size: icosphere instances (each its
    own occluder with its own bounding box) on the terrain, a 3840x2160
    street-level view (first-block wrap included) and two overhead 1024^2
    views in one step, each against the oracle's classify-then-draw."""
    sizes = [(3840, 2160), (1024, 1024), (1024, 1024)]
    sc = scenes.config5(n_instances=2000, terrain_n=128, queries_per_view=2048, sizes=sizes)
    views = sc.meta["views"]
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    setup = []
    for k, vw in enumerate(views):
        occ = [sc.occluders[i] for i in scenes.front_to_back(sc.occluders, vw["eye"])]
        V.set_mvp(k, vw["mvp"])
        V.set_occluders(k, occ)
        V.set_queries(k, vw["queries"])
        setup.append(occ)
    V.run(graph=True)
    V.sync()
    drawn = 0
    for k, vw in enumerate(views):
        w, h = vw["size"]
        drawn += _check_view(oracle_mod, V, k, w, h, vw["mvp"], setup[k], vw["queries"], report,
                             f"stress view {k} {w}x{h}")
    assert drawn > 500

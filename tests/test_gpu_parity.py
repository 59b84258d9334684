"""GPU parity: the HIP path against the CPU oracle on identical inputs.

Raw depth, hiZ, readBackDepth bytes, draw statuses and query bits must be
bit-exact; any depth difference is reported as a count and a max LSB."""

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, scenes

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, w, h, mvp):
    g = Rasterizer(w, h)
    o = oracle_mod.OracleRasterizer(w, h)
    for r in (g, o):
        r.clear()
        r.setModelViewProjection(mvp)
    return g, o


def _assert_same_buffers(g, o, report, what):
    dg, hg = g.raw()
    do, ho = o.raw()
    assert np.array_equal(hg, ho), f"{what}: hiZ {report(hg, ho)}"
    assert np.array_equal(dg, do), f"{what}: depth {report(dg, do)}"


@pytest.fixture(scope="module")
def scene1():
    return scenes.config1(seed=1)


def test_config1_frame(oracle_mod, scene1, report):
    m = scene1.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvp"])
    order = scenes.front_to_back(scene1.occluders, m["eye"])
    occ = [scene1.occluders[i] for i in order]
    print(f"config1 seed 1: {len(occ)} occluders, {scene1.quad_count} quads", flush=True)
    st_o = o.draw_frame(occ)
    st_g = g.draw_frame(occ)
    assert np.array_equal(st_g, st_o), f"statuses gpu {st_g.tolist()} oracle {st_o.tolist()}"
    _assert_same_buffers(g, o, report, "config1 frame")
    assert np.array_equal(g.readBackDepth(), o.readBackDepth())
    vg, cg = g.query_batch(m["queries"])
    vo, co = o.query_batch(m["queries"])
    assert np.array_equal(vg, vo), f"{int((vg != vo).sum())} query bits differ"
    assert np.array_equal(cg, co)


def test_many_draws_without_sync(oracle_mod, scene1, report):
    """Forty query-gated draws into one buffer with no sync in between, each
    of a random run of the front-to-back order (1 to all occluders): the
    host's item-table slots are reused, and regrown, while earlier draws are
    still in flight.  The oracle draws the same sequence."""
    m = scene1.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvp"])
    order = scenes.front_to_back(scene1.occluders, m["eye"])
    occ = [scene1.occluders[i] for i in order]
    rng = np.random.default_rng(17)
    sizes = []
    for _ in range(40):
        n = int(rng.integers(1, len(occ) + 1))
        start = int(rng.integers(0, len(occ) - n + 1))
        run = occ[start:start + n]
        g.draw_frame_handles(g.handle_array(run), n)
        o.draw_frame(run)
        sizes.append(n)
    print(f"draw sizes {sizes}", flush=True)
    _assert_same_buffers(g, o, report, "forty draws")


def test_config1_dropin_calls(oracle_mod, scene1, report):
    """Main.cpp:192-206 through the per-call API: queryVisibility (blocking),
    then rasterize<needsClipping> (enqueued)."""
    m = scene1.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvp"])
    order = scenes.front_to_back(scene1.occluders, m["eye"])
    for i in order:
        occ = scene1.occluders[i]
        vg, cg = g.queryVisibility(occ.m_boundsMin, occ.m_boundsMax)
        vo, co = o.queryVisibility(occ.m_boundsMin, occ.m_boundsMax)
        assert (vg, cg) == (vo, co), f"occluder {i}"
        if vo:
            g.rasterize(occ, co)
            o.rasterize(occ, co)
    _assert_same_buffers(g, o, report, "config1 per-call")


@pytest.mark.parametrize("clip", [False, True])
def test_config1_rasterize_all(oracle_mod, scene1, report, clip):
    """rasterize<clip> of every occluder, no query gating (clip=False also
    exercises behind-camera quads on the unclipped path)."""
    m = scene1.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvp"])
    for occ in scene1.occluders:
        g.rasterize(occ, clip)
        o.rasterize(occ, clip)
    _assert_same_buffers(g, o, report, f"rasterize<{clip}> all")


@pytest.fixture(scope="module")
def scene2():
    return scenes.config2(seed=1)


def test_config2_camera_path(oracle_mod, scene2, report):
    """All 256 frames of the city path on one reused buffer (stale depth
    between frames, as the reference's clear leaves it), 1920x1080."""
    m = scene2.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvps"][0])
    nf = m["mvps"].shape[0]
    drawn = clipped = 0
    for f in range(nf):
        for r in (g, o):
            r.clear()
            r.setModelViewProjection(m["mvps"][f])
        occ = [scene2.occluders[i] for i in scenes.front_to_back(scene2.occluders, m["eyes"][f])]
        st_o = o.draw_frame(occ)
        st_g = g.draw_frame(occ)
        assert np.array_equal(st_g, st_o), f"frame {f}: statuses differ at {np.nonzero(st_g != st_o)[0][:10]}"
        drawn += int((st_o > 0).sum())
        clipped += int((st_o == 2).sum())
        _assert_same_buffers(g, o, report, f"config2 frame {f}")
        if f % 64 == 63:
            vg, cg = g.query_batch(m["queries"])
            vo, co = o.query_batch(m["queries"])
            assert np.array_equal(vg, vo), f"frame {f}: {int((vg != vo).sum())} query bits differ"
            assert np.array_equal(cg, co), f"frame {f}: needsClipping bits differ"
            assert np.array_equal(g.readBackDepth(), o.readBackDepth()), f"frame {f}: readback differs"
    print(f"config2: {nf} frames, {drawn} occluder draws, {clipped} with rasterize<true>", flush=True)
    assert clipped > 0, "the path should exercise the near-clipped raster path"


def test_counting_variant_matches_oracle(oracle_mod, scene1, report):
    """The observability build of the frame kernel (per-tile counters) draws
    the same bits as the oracle, and its counters are consistent: every
    visited (quad, block) pair is either rejected or written, early query
    answers are a subset of the early evaluations, and every tile's total
    covers its item loop."""
    m = scene1.meta
    g, o = _pair(oracle_mod, m["width"], m["height"], m["mvp"])
    occ = [scene1.occluders[i] for i in scenes.front_to_back(scene1.occluders, m["eye"])]
    g.frame_stats(True)
    st_g = g.draw_frame(occ)
    st_o = o.draw_frame(occ)
    assert np.array_equal(st_g, st_o)
    _assert_same_buffers(g, o, report, "counting variant")
    s = g.frame_stats_read().astype(np.int64)
    tw = g.config()["tileW"]
    assert s.shape == (-(-g.blocks_x // tw) * -(-g.blocks_y // tw), 22)
    assert np.array_equal(s[:, 13], s[:, 14] + s[:, 15])
    assert np.all(s[:, 21] <= s[:, 20])
    assert np.all(s[:, 0] >= s[:, 11]) and s[:, 6].sum() > 0
    print(f"visited {s[:, 13].sum()}, written {s[:, 15].sum()}, early tried {s[:, 20].sum()} "
          f"answered {s[:, 21].sum()}", flush=True)
    g.frame_stats(False)


def test_crowded_tile_item_windows(oracle_mod, report):
    """Hundreds of small occluders crowded into one or two screen tiles, so a
    tile's item windows overflow: more than 128 of a window's candidates are
    relevant, the first 128 are kept and the next window starts at the first
    one not kept.  The single-view frame (all items as candidates), gated
    views (binned lists, every candidate relevant) and classify-then-draw
    views, against the oracle bit for bit."""
    from rasterizer_amd import Views

    rng = scenes.SplitMix64(21)
    n = 1200
    c, h = scenes.random_boxes(rng, n, -2.5, 2.5, 0.05, 0.3)
    c[:, 2] += 100.0  # a cluster about 40 px across at 1280x720
    bg_c, bg_h = scenes.random_boxes(rng, 200, -50.0, 50.0, 0.5, 5.0)
    bg_c[:, 2] = np.abs(bg_c[:, 2]) + 60.0
    quads = np.concatenate([scenes.box_quads(c[i], h[i]) for i in range(n)] +
                           [scenes.box_quads(bg_c[i], bg_h[i]) for i in range(200)], axis=0)
    scene = scenes.bake_scene(quads, target=24, gran=8)  # one to three batches per occluder
    w, h = 1280, 720
    eye = np.zeros(3, np.float32)
    mvp = scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), np.pi / 3, w / h)
    occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
    qrng = scenes.SplitMix64(22)
    qc, qh = scenes.random_boxes(qrng, 2048, -3.0, 3.0, 0.02, 0.2)
    qc[:, 2] += 101.0
    queries = scenes.aabb_bounds(qc, qh)
    print(f"crowded tile: {len(occ)} occluders, {scene.quad_count} quads", flush=True)
    g, o = _pair(oracle_mod, w, h, mvp)
    st_o = o.draw_frame(occ)
    st_g = g.draw_frame(occ)
    assert np.array_equal(st_g, st_o), f"statuses differ at {np.nonzero(st_g != st_o)[0][:10]}"
    _assert_same_buffers(g, o, report, "crowded tile frame")
    vg, cg = g.query_batch(queries)
    vo, co = o.query_batch(queries)
    assert np.array_equal(vg, vo) and np.array_equal(cg, co)
    # The cluster (|x|, |y| < 3 at z ~ 100) spans about 34 x 34 px around the
    # screen centre, so its occluders all touch tiles (9, 5) and (10, 5): far
    # more relevant items than one window keeps.
    ctr = np.array([(o.m_boundsMin[:3] + o.m_boundsMax[:3]) / 2 for o in occ])
    assert int(((np.abs(ctr[:, 0]) < 3) & (np.abs(ctr[:, 1]) < 3) & (np.abs(ctr[:, 2] - 100) < 3)).sum()) > 300
    for gated in (True, False):
        V = Views(Rasterizer(64, 64), [(w, h)])
        V.set_query_gated(gated)
        V.set_mvp(0, mvp)
        V.set_occluders(0, occ)
        V.set_queries(0, queries)
        V.run(graph=False)
        V.sync()
        ov = oracle_mod.OracleRasterizer(w, h)
        ov.clear()
        ov.setModelViewProjection(mvp)
        st_v = ov.draw_frame(occ, gated=gated)
        assert np.array_equal(V.status(0), st_v), f"views gated={gated}: statuses differ"
        dg, hg = V.raw(0)
        do, ho = ov.raw()
        assert np.array_equal(hg, ho), f"views gated={gated}: hiZ {report(hg, ho)}"
        assert np.array_equal(dg, do), f"views gated={gated}: depth {report(dg, do)}"
        vq, cq = V.queries(0)
        oq, ocq = ov.query_batch(queries)
        assert np.array_equal(vq, oq) and np.array_equal(cq, ocq)

"""Launch configurations and the autotuner (rz_set_config, rz_tune_frame,
rz_views_tune): every configuration gives the oracle's bits; the tuner keeps
the fastest candidate and records it in a cache file that later contexts
load, explicitly or through RZ_TUNE_CACHE."""

from __future__ import annotations

import os

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, RasterizerError, Views, scenes

pytestmark = pytest.mark.gpu

CONFIGS = [(8, 2), (8, 4), (8, 8), (16, 2), (16, 4), (16, 8)]


@pytest.fixture(scope="module")
def scene():
    return scenes.config1(seed=3, n_boxes=768)


def _oracle_frame(oracle_mod, scene, w, h):
    o = oracle_mod.OracleRasterizer(w, h)
    o.clear()
    o.setModelViewProjection(scene.meta["mvp"])
    st = o.draw_frame(scene.occluders, gated=True)
    return st, o.raw()


@pytest.mark.parametrize("tile_w,waves", CONFIGS)
def test_every_frame_config_matches_oracle(oracle_mod, scene, report, tile_w, waves):
    w, h = scene.meta["width"], scene.meta["height"]
    R = Rasterizer(w, h)
    R.set_config(tile_w, waves)
    assert R.config() == {"tileW": tile_w, "waves": waves}
    R.setModelViewProjection(scene.meta["mvp"])
    st_o, (do, ho) = _oracle_frame(oracle_mod, scene, w, h)
    for _ in range(2):  # the second frame starts from the first one's buffer
        R.clear()
        st = R.draw_frame(scene.occluders)
        assert np.array_equal(st, st_o)
        d, hz = R.raw()
        assert np.array_equal(hz, ho), report(hz, ho)
        assert np.array_equal(d, do), report(d, do)


def test_frame_tuner_picks_fastest_and_persists(oracle_mod, scene, report, tmp_path, monkeypatch):
    w, h = scene.meta["width"], scene.meta["height"]
    cache = tmp_path / "tune.txt"
    R = Rasterizer(w, h)
    R.setModelViewProjection(scene.meta["mvp"])
    res = R.autotune(scene.occluders, reps=5, cache=cache)
    cands = res["candidates"]
    print(res, flush=True)
    assert [(c["tileW"], c["waves"]) for c in cands] == CONFIGS
    timed = [c for c in cands if c["ms"] > 0]
    assert len(timed) == len(CONFIGS)  # 1280x720 fits every configuration
    best = min(timed, key=lambda c: c["ms"])
    assert res["chosen"] == {"tileW": best["tileW"], "waves": best["waves"]}
    # The tuner leaves the frame drawn, with the oracle's bits.
    _, (do, ho) = _oracle_frame(oracle_mod, scene, w, h)
    d, hz = R.raw()
    assert np.array_equal(hz, ho) and np.array_equal(d, do), report(d, do)
    text = cache.read_text()
    assert text.count("\n") == 1 and f"frame,{w}x{h}\t" in text
    # Another size never matches the entry; this size does.
    other = Rasterizer(640, 360)
    assert not other.load_config(cache)
    R2 = Rasterizer(w, h)
    R2.set_config(8 if best["tileW"] == 16 else 16, 4)
    assert R2.load_config(cache)
    assert R2.config() == res["chosen"]
    # Contexts created with RZ_TUNE_CACHE load it themselves.
    monkeypatch.setenv("RZ_TUNE_CACHE", os.fspath(cache))
    R3 = Rasterizer(w, h)
    assert R3.config() == res["chosen"]
    # Re-tuning replaces the entry instead of adding one.
    R.autotune(scene.occluders, reps=3, cache=cache)
    assert cache.read_text().count("\n") == 1


def test_config_rejects_bad_and_too_large():
    R = Rasterizer(3840, 2160)
    before = R.config()
    with pytest.raises(RasterizerError):
        R.set_config(8, 5)
    with pytest.raises(RasterizerError):
        R.set_config(12, 4)
    # 60 x 34 tiles of 8x8 blocks cannot all be co-resident on one MI355X.
    with pytest.raises(RasterizerError):
        R.set_config(8, 8)
    assert R.config() == before


def test_views_tuner(oracle_mod, scene, report, tmp_path):
    sizes = [(320, 192), (640, 360), (128, 64), (1280, 720)]
    V = Views(Rasterizer(64, 64), sizes)
    rng = np.random.default_rng(8)
    setup = []
    for k, (w, h) in enumerate(sizes):
        eye = rng.uniform(-40, 40, 3).astype(np.float32)
        mvp = scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), np.pi / 3, w / h)
        occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
        q = scene.meta["queries"][:500]
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
        V.set_queries(k, q)
        setup.append((mvp, occ, q))
    V.enable_labels(True)
    cache = tmp_path / "views.txt"
    res = V.autotune(reps=3, cache=cache)
    print(res, flush=True)
    cands = [c for c in res["candidates"] if c["ms"] > 0]
    assert len(cands) >= 3
    best = min(cands, key=lambda c: c["ms"])
    assert res["chosen"] == {"tileW": best["tileW"], "waves": best["waves"], "groupsPerCU": best["groupsPerCU"]}
    # Labelled steps draw with 8x8-block tiles: the 16x16 candidates are not timed.
    assert all(c["ms"] < 0 for c in res["candidates"] if c["tileW"] == 16)
    V.run(graph=True)
    V.sync()
    for k, (w, h) in enumerate(sizes):
        mvp, occ, q = setup[k]
        o = oracle_mod.OracleRasterizer(w, h)
        o.clear()
        o.setModelViewProjection(mvp)
        st, lab = o.frame_labels(occ)
        assert np.array_equal(V.status(k), st)
        dg, hg = V.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hg, ho), report(hg, ho)
        written = np.repeat(ho[: (w // 8) * (h // 8)] != 1, 64)
        assert np.array_equal(dg[written], do[written]), report(dg[written], do[written])
        vg, cg = V.queries(k)
        vo, co = o.query_batch(q)
        assert np.array_equal(vg, vo) and np.array_equal(cg, co)
        assert np.array_equal(V.labels(k)[0], lab)
    V2 = Views(Rasterizer(64, 64), sizes)
    assert V2.load_config(cache) and V2.config() == res["chosen"]


def test_views_tuner_tile_sizes(oracle_mod, scene, report):
    """Without labels the tuner also times 16x16-block (128-pixel) tiles; every
    candidate gives the same bits, and the chosen configuration's frames equal
    the oracle's."""
    sizes = [(1280, 720), (512, 512), (200, 104)]
    V = Views(Rasterizer(64, 64), sizes)
    rng = np.random.default_rng(18)
    setup = []
    for k, (w, h) in enumerate(sizes):
        eye = rng.uniform(-40, 40, 3).astype(np.float32)
        mvp = scenes.view_projection(eye, (0, 0, 1), (0, 1, 0), np.pi / 3, w / h)
        occ = [scene.occluders[i] for i in scenes.front_to_back(scene.occluders, eye)]
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
        V.set_queries(k, scene.meta["queries"][:300])
        setup.append((mvp, occ))
    res = V.autotune(reps=2)
    print(res, flush=True)
    timed = {(c["tileW"], c["waves"], c["groupsPerCU"]) for c in res["candidates"] if c["ms"] > 0}
    assert any(t[0] == 16 for t in timed) and any(t[0] == 8 for t in timed)
    for tile_w in (8, 16):
        V.set_config(8, 0, tile_w)
        V.run(graph=True)
        V.sync()
        for k, (w, h) in enumerate(sizes):
            mvp, occ = setup[k]
            o = oracle_mod.OracleRasterizer(w, h)
            o.clear()
            o.setModelViewProjection(mvp)
            assert np.array_equal(V.status(k), o.draw_frame(occ, gated=False))
            dg, hg = V.raw(k)
            do, ho = o.raw()
            written = np.repeat(ho[: (w // 8) * (h // 8)] != 1, 64)
            assert np.array_equal(hg, ho), f"tile {tile_w} view {k}: hiZ {report(hg, ho)}"
            assert np.array_equal(dg[written], do[written]), f"tile {tile_w} view {k}: depth"

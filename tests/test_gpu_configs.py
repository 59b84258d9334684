"""GPU parity at the full sizes of the multi-view configurations, every view
against the oracle bit for bit (statuses, raw depth + hiZ, query bits).

* Batched multi-view (scenes.config3): 64 views per step (16 x 2048^2,
  24 x 1024^2, 16 x 512^2, 8 x 1920x1080), per-view occluder subsets of 1k to
  200k quads from a 40,000-box scene, 16,384 queries per view; two steps of
  graph replays, the second from the buffers the first left (clear leaves
  depth stale).
* Learned draw order (config 1's 1,024-box scene at 1280x720 with the shipped
  model): the order the model picks, against the oracle drawing the same
  batches in the same order; the buffer equals the front-to-back buffer, and
  the hiZ early-reject rates of both orders are reported.
* Stress (scenes.config5): 50,000 icosphere instances on the terrain, 4.26M
  quads, one 3840x2160 and fifteen 1024^2 views each submitting all of them
  front to back, 65,536 queries per view (1M in total)."""

from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, Views, learn, scenes

pytestmark = pytest.mark.gpu


def _oracles(oracle_mod, sizes):
    return {s: oracle_mod.OracleRasterizer(*s) for s in sorted(set(sizes))}


def _check(o, V, k, mvp, occ, q, before, report, what):
    """View k of V against oracle o run from the buffer `before` (depth, hiZ)."""
    o.set_raw(*before)
    o.clear()
    o.setModelViewProjection(mvp)
    st = o.draw_frame(occ, gated=False)
    assert np.array_equal(V.status(k), st), f"{what}: statuses differ at {np.flatnonzero(V.status(k) != st)[:10]}"
    dg, hg = V.raw(k)
    do, ho = o.raw()
    assert np.array_equal(hg, ho), f"{what}: hiZ {report(hg, ho)}"
    assert np.array_equal(dg, do), f"{what}: depth {report(dg, do)}"
    if q is not None:
        vg, cg = V.queries(k)
        vo, co = o.query_batch(q)
        assert np.array_equal(vg, vo), f"{what}: {int((vg != vo).sum())} query bits differ"
        assert np.array_equal(cg, co), f"{what}: {int((cg != co).sum())} clip bits differ"
    return dg, hg, int((st > 0).sum())


def test_config3_full(oracle_mod, report):
    sc = scenes.config3()
    views = sc.meta["views"]
    sizes = [v["size"] for v in views]
    print(f"config3: {len(views)} views, {sum(v['quads'] for v in views):,} quads, "
          f"{sum(v['queries'].shape[0] for v in views):,} queries per step", flush=True)
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    for k, v in enumerate(views):
        V.set_occluders(k, v["occluders"])
        V.set_queries(k, v["queries"])
    oracles = _oracles(oracle_mod, sizes)
    before = [(np.zeros((w // 8) * (h // 8) * 64, np.uint16), np.zeros((w // 8) * (h // 8) + 8, np.uint16))
              for w, h in sizes]
    for step in range(2):
        for k, v in enumerate(views):
            V.set_mvp(k, scenes.view_mvp(v, step))
        V.run(graph=True)
        V.sync()
        drawn = 0
        for k, v in enumerate(views):
            w, h = v["size"]
            d, hz, n = _check(oracles[(w, h)], V, k, scenes.view_mvp(v, step), v["occluders"], v["queries"],
                              before[k], report, f"step {step} view {k} {w}x{h}")
            before[k] = (d, hz)
            drawn += n
        print(f"config3 step {step}: {drawn} occluders drawn, every view bit-exact", flush=True)


def _batch_parts(occ):
    return [(i, b) for i, o in enumerate(occ) for b in range(o.packets.size // 32)]


def test_config4_learned_order_full(oracle_mod, report):
    """The shipped model orders every batch of the 1,024-box scene for 16
    cameras at 1280x720 (eyes in the scene's cube, random directions, as the
    model was trained).  Step 1's features carry step 0's labels."""
    sc = scenes.config1(seed=1)
    w, h = 1280, 720
    nv = 16
    rng = np.random.default_rng(44)
    cams = [learn.random_camera(rng) for _ in range(nv)]
    owner = Rasterizer(64, 64)
    VL = Views(owner, [(w, h)] * nv)
    VF = Views(owner, [(w, h)] * nv)
    setups = []
    for k, (eye, d, up) in enumerate(cams):
        occ = [sc.occluders[i] for i in scenes.front_to_back(sc.occluders, eye)]
        setups.append((eye, d, up, occ))
        for V in (VL, VF):
            V.set_occluders(k, occ)
    VL.set_model(learn.load(), by_occluder=False)
    for V in (VL, VF):
        V.enable_stats(True)
    before = None
    for step in range(2):
        if step == 1:
            before = [VL.raw(k) for k in range(nv)]  # what step 0 left in the learned-order buffers
        for k, (eye, d, up, _) in enumerate(setups):
            mvp = scenes.view_projection(eye - (0.5 if step == 0 else 0.0) * d, d, up, math.pi / 3, w / h)
            for V in (VL, VF):
                V.set_mvp(k, mvp)
        for V in (VL, VF):
            V.run(graph=True)
            V.sync()
    sl, sf = VL.stats(), VF.stats()
    mismatch = maxlsb = 0
    o = oracle_mod.OracleRasterizer(w, h)
    for k, (eye, d, up, occ) in enumerate(setups):
        mvp = scenes.view_projection(eye, d, up, math.pi / 3, w / h)
        order = VL.order(k)
        parts = _batch_parts(occ)
        assert sorted(order.tolist()) == list(range(len(parts)))
        # The oracle draws the same batches in the model's order (each with its
        # occluder's classification), from the buffer step 0 left.
        o.set_raw(*before[k])
        o.clear()
        o.setModelViewProjection(mvp)
        st = o.draw_frame(occ, gated=False)
        o.set_raw(*before[k])
        o.clear()
        for j in order:
            i, b = parts[j]
            if st[i]:
                src = occ[i]
                o.rasterize(SimpleNamespace(packets=src.packets[32 * b:32 * b + 32], m_refMin=src.m_refMin,
                                            m_refMax=src.m_refMax), st[i] == 2)
        dl, hl = VL.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hl, ho), f"view {k}: learned order hiZ {report(hl, ho)}"
        assert np.array_equal(dl, do), f"view {k}: learned order depth {report(dl, do)}"
        df, hf = VF.raw(k)
        diff = np.abs(dl.astype(np.int64) - df.astype(np.int64))
        mismatch += int((diff > 0).sum()) + int((hl != hf).sum())
        maxlsb = max(maxlsb, int(diff.max()))
    rl = sl["rejected"] / max(sl["visited"], 1)
    rf = sf["rejected"] / max(sf["visited"], 1)
    print(f"config4: learned vs front-to-back buffers: {mismatch} values differ, max {maxlsb} LSB; "
          f"hiZ/edge reject rate learned {rl:.4f}, front-to-back {rf:.4f}", flush=True)
    # The merge is a per-pixel max and a cleared block's first write covers the
    # whole block, so the final buffer does not depend on the draw order.
    assert mismatch == 0


def test_config5_full(oracle_mod, report):
    sc = scenes.config5()
    views = sc.meta["views"]
    sizes = [v["size"] for v in views]
    nq = sum(v["queries"].shape[0] for v in views)
    print(f"config5: {len(sc.occluders):,} occluders, {sc.quad_count:,} quads, {len(views)} views, "
          f"{nq:,} queries", flush=True)
    owner = Rasterizer(64, 64)
    V = Views(owner, sizes)
    setup = []
    for k, v in enumerate(views):
        occ = [sc.occluders[i] for i in scenes.front_to_back(sc.occluders, v["eye"])]
        V.set_mvp(k, v["mvp"])
        V.set_occluders(k, occ)
        V.set_queries(k, v["queries"])
        setup.append(occ)
    V.run(graph=True)
    V.sync()
    oracles = _oracles(oracle_mod, sizes)
    drawn = 0
    for k, v in enumerate(views):
        w, h = v["size"]
        zero = (np.zeros((w // 8) * (h // 8) * 64, np.uint16), np.zeros((w // 8) * (h // 8) + 8, np.uint16))
        drawn += _check(oracles[(w, h)], V, k, v["mvp"], setup[k], v["queries"], zero, report,
                        f"view {k} {w}x{h}")[2]
        print(f"config5 view {k} {w}x{h}: bit-exact", flush=True)
    assert drawn > 1000

"""GPU parity on known-answer scenes and edge cases, each against the oracle
bit for bit: host tables, every readBackDepth value, query2D on random
buffers, query boxes around the frustum and the eye, the pinned
known-answer scenes, the 16-bit first-block wrap at 3840x2160, tiny buffers
and empty frames."""

from __future__ import annotations

import math

import numpy as np
import pytest

import kat
from rasterizer_amd import Occluder, Rasterizer, scenes

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, w, h, mvp=None):
    if mvp is None:
        mvp, _ = kat.camera(w, h)
    g = Rasterizer(w, h)
    o = oracle_mod.OracleRasterizer(w, h)
    for r in (g, o):
        r.clear()
        r.setModelViewProjection(mvp)
    return g, o


def _same(g, o, report, what):
    dg, hg = g.raw()
    do, ho = o.raw()
    assert np.array_equal(hg, ho), f"{what}: hiZ {report(hg, ho)}"
    assert np.array_equal(dg, do), f"{what}: depth {report(dg, do)}"


def _product_occluder(oracle_mod, quads):
    """Bakes with the oracle (so both sides draw identical packets) and wraps
    the result as the product's Occluder."""
    occ = kat.occluder_from_quads(oracle_mod.bake, quads)
    return occ, Occluder(occ.packets, occ.m_refMin, occ.m_refMax, occ.m_boundsMin, occ.m_boundsMax, occ.m_center)


def test_tables_match_oracle(oracle_mod):
    """The coverage table built at context creation equals the oracle's, and
    the RCPPS table holds the host's rcp of 1.m for each top-12-bit mantissa."""
    g = Rasterizer(64, 64)
    cov, rcp = g.tables()
    assert np.array_equal(cov, oracle_mod.OracleRasterizer(64, 64).table())
    for i in range(0, 4096, 37):
        x = np.array([0x3F800000 | (i << 11)], np.uint32).view(np.float32)[0]
        want = np.array([oracle_mod.x86_rcp(float(x))], np.float32).view(np.uint32)[0]
        assert rcp[i] == want, i


def test_set_mvp_state_matches(oracle_mod):
    mvp, _ = kat.camera(320, 200)
    g, o = _pair(oracle_mod, 320, 200, mvp)
    pg, rg = g.mvp_state()
    po, ro = o.mvp_state()
    assert np.array_equal(pg, po) and np.array_equal(rg, ro)


@pytest.mark.parametrize("size", [(72, 40), (8, 8), (16, 8)])
def test_clear_padding(oracle_mod, size, report):
    """clear on odd block counts: whole 8-entry groups, padding kept."""
    w, h = size
    g, o = _pair(oracle_mod, w, h)
    nb = (w // 8) * (h // 8)
    rng = np.random.default_rng(nb)
    depth = rng.integers(0, 65536, nb * 64, dtype=np.uint16)
    hiz = np.full(nb + 8, 777, np.uint16)
    for r in (g, o):
        r.upload_raw(depth, hiz) if r is g else r.set_raw(depth, hiz)
        r.clear()
    _same(g, o, report, f"clear {w}x{h}")


def test_readback_every_value(oracle_mod):
    w = h = 256
    img = np.arange(65536, dtype=np.uint32).astype(np.uint16).reshape(h, w)
    hiz = np.zeros((w // 8) * (h // 8) + 8, np.uint16)
    hiz[5] = 1  # one fast-cleared block reads back as zeros
    g, o = _pair(oracle_mod, w, h)
    g.upload_raw(kat.pack_block_linear(img), hiz)
    o.set_raw(kat.pack_block_linear(img), hiz)
    assert np.array_equal(g.readBackDepth(), o.readBackDepth())


def test_query2d_random_buffers(oracle_mod):
    w, h = 128, 64
    g, o = _pair(oracle_mod, w, h)
    rng = np.random.default_rng(11)
    for trial in range(3):
        img = rng.integers(30000, 40000, (h, w)).astype(np.uint16)
        blocks = img.reshape(h // 8, 8, w // 8, 8).min(axis=(1, 3))
        sel = rng.random(blocks.shape)
        blocks[sel < 0.15] = 1
        blocks[sel > 0.9] = 45000
        hiz = np.concatenate([blocks.reshape(-1), np.zeros(8, np.uint16)]).astype(np.uint16)
        g.upload_raw(kat.pack_block_linear(img), hiz)
        o.set_raw(kat.pack_block_linear(img), hiz)
        for _ in range(150):
            x0, x1 = sorted(int(v) for v in rng.integers(0, w, 2))
            y0, y1 = sorted(int(v) for v in rng.integers(0, h, 2))
            z = int(rng.integers(28000, 46000))
            assert g.query2D(x0, x1, y0, y1, z) == o.query2D(x0, x1, y0, y1, z), (trial, x0, x1, y0, y1, z)


def _edge_boxes(rng, n):
    """Boxes behind, around and in front of the eye, off-screen, flat and
    degenerate (min == max), tiny and huge."""
    out = []
    for _ in range(n):
        c = rng.uniform(-60, 60, 3) * np.array([1, 1, 1.5])
        e = np.exp(rng.uniform(math.log(0.001), math.log(50), 3))
        kind = rng.integers(0, 6)
        if kind == 0:
            e[rng.integers(0, 3)] = 0.0
        elif kind == 1:
            e[:] = 0.0
        elif kind == 2:
            c[2] = rng.uniform(-2, 2)
        out.append(np.concatenate([c - e, [1.0], c + e, [1.0]]))
    return np.array(out, np.float32)


def test_query_batch_edge_boxes(oracle_mod):
    """queryVisibility over boxes around the eye and frustum edges, against
    an empty buffer and against a drawn scene."""
    sc = scenes.config1(seed=4, n_boxes=256, width=320, height=192)
    m = sc.meta
    g, o = _pair(oracle_mod, 320, 192, m["mvp"])
    boxes = _edge_boxes(np.random.default_rng(2), 3000)
    for phase in range(2):
        vg, cg = g.query_batch(boxes)
        vo, co = o.query_batch(boxes)
        assert np.array_equal(vg, vo), f"phase {phase}: {int((vg != vo).sum())} visibility bits differ"
        assert np.array_equal(cg, co), f"phase {phase}: {int((cg != co).sum())} clip bits differ"
        occ = [sc.occluders[i] for i in scenes.front_to_back(sc.occluders, m["eye"])]
        g.draw_frame(occ)
        o.draw_frame(occ)


def _kat_scenes():
    """(name, width, height, [(quads, clip), ...]) of the oracle's known-answer scenes."""
    out = []
    for i, (rect, dist) in enumerate([((-3.0, 2.0, -1.5, 2.5), 10.0), ((-40.0, 35.0, -22.0, 30.0), 100.0),
                                      ((-0.9, 0.7, -0.3, 0.6), 2.0), ((-2000.0, 2000.0, -1000.0, 1000.0), 500.0)]):
        for clip in (False, True):
            out.append((f"rect{i}-{clip}", 256, 144, [([kat.facing_rect(*rect, dist)], clip)]))
    w, h = 128, 96
    ux = 10.0 * math.tan(math.pi / 6) * (w / h) / (w * 0.5 - 4.0)
    uy = 10.0 * math.tan(math.pi / 6) / (h * 0.5 - 4.0)
    for shift in range(8):
        x0 = (-20 + shift) * ux + 0.1 * ux
        y0 = (-12 + shift) * uy + 0.1 * uy
        q = kat.facing_rect(x0, x0 + (13 + shift) * ux, y0, y0 + (11 + shift) * uy, 10.0)
        out.append((f"residue{shift}", w, h, [([q], False)]))
    c = np.array([0.5, -0.3, 12.0])
    u = np.array([3.0, 0.0, 1.5])
    v = np.array([0.0, 2.0, -0.8])
    slanted = np.array([c - u - v, c + u - v, c + u + v, c - u + v])
    for clip in (False, True):
        out.append((f"slanted-{clip}", 192, 128, [([slanted], clip)]))
    y = -2.0
    floor = np.array([[-30.0, y, -20.0], [30.0, y, -20.0], [30.0, y, 300.0], [-30.0, y, 300.0]])
    wall = kat.facing_rect(-400.0, 400.0, -300.0, 300.0, 100.0)
    out.append(("floor", 192, 128, [([floor], True)]))
    out.append(("floor-over-wall", 192, 128, [([wall], False), ([floor], True)]))
    out.append(("floor-unclipped", 192, 128, [([floor], False)]))  # behind-eye quad on the <false> path
    near = kat.facing_rect(-2.0, 1.0, -1.0, 1.0, 8.0)
    far = kat.facing_rect(-1.0, 3.0, -2.0, 0.5, 20.0)
    out.append(("merge", 128, 96, [([near], False), ([far], False)]))
    out.append(("backface", 128, 64, [([kat.facing_rect(-3.0, 2.0, -1.5, 2.5, 10.0)[::-1]], True)]))
    return out


@pytest.mark.parametrize("scene", _kat_scenes(), ids=lambda s: s[0])
def test_known_answer_scenes(oracle_mod, scene, report):
    name, w, h, draws = scene
    print(f"scene {name} {w}x{h}", flush=True)
    g, o = _pair(oracle_mod, w, h)
    for quads, clip in draws:
        occ_o, occ_g = _product_occluder(oracle_mod, quads)
        o.rasterize(occ_o, clip)
        g.rasterize(occ_g, clip)
    _same(g, o, report, name)
    assert np.array_equal(g.readBackDepth(), o.readBackDepth())


def test_first_block_wrap_4k(oracle_mod, report):
    """3840x2160 has 129,600 blocks: the reference's 16-bit multiply of the
    first block index (Rasterizer.cpp:1054) wraps for quads starting at block
    row 137 or below, and the wrapped index is what it draws to."""
    sc = scenes.config1(seed=5, n_boxes=384, width=3840, height=2160, n_queries=512)
    m = sc.meta
    g, o = _pair(oracle_mod, 3840, 2160, m["mvp"])
    print(f"4K: {len(sc.occluders)} occluders", flush=True)
    for clip in (False, True):
        for occ in sc.occluders:
            o.rasterize(occ, clip)
            g.rasterize(occ, clip)
        _same(g, o, report, f"4K rasterize<{clip}>")
    vg, cg = g.query_batch(m["queries"])
    vo, co = o.query_batch(m["queries"])
    assert np.array_equal(vg, vo) and np.array_equal(cg, co)


@pytest.mark.parametrize("size", [(8, 8), (16, 8), (8, 24), (40, 16)])
def test_tiny_buffers(oracle_mod, size, report):
    w, h = size
    sc = scenes.config1(seed=6, n_boxes=128, width=w, height=h, n_queries=256)
    m = sc.meta
    g, o = _pair(oracle_mod, w, h, m["mvp"])
    occ = [sc.occluders[i] for i in scenes.front_to_back(sc.occluders, m["eye"])]
    st_g = g.draw_frame(occ)
    st_o = o.draw_frame(occ)
    assert np.array_equal(st_g, st_o)
    _same(g, o, report, f"{w}x{h}")
    vg, cg = g.query_batch(m["queries"])
    vo, co = o.query_batch(m["queries"])
    assert np.array_equal(vg, vo) and np.array_equal(cg, co)


def test_empty_frame_and_occluder(oracle_mod, report):
    g, o = _pair(oracle_mod, 64, 64)
    assert g.draw_frame([]).size == 0
    g.sync()
    _same(g, o, report, "empty frame")
    with pytest.raises(ValueError):
        Occluder(np.zeros(8, np.uint32), np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))


def test_degenerate_reference_box_matches_oracle(oracle_mod, report):
    """A flat occluder facing an axis-aligned camera, baked against a box 1e-5
    deep: the prologue's c0 is 0/0 (Rasterizer.cpp:645-655).  x86 gives the
    negative default NaN, every depth is that NaN and packs to maxZ 0, so no
    block is written; the GPU must follow x86's NaN rules, not its own."""
    from rasterizer_amd import Views
    from test_oracle_pins import _flat_facing_occluder

    occ = _flat_facing_occluder()
    mvp = scenes.view_projection((0, 0, 0), (0, 0, 1), (0, 1, 0), np.pi / 3, 2.0)
    g, o = _pair(oracle_mod, 128, 64, mvp)
    for clip in (False, True):
        g.rasterize(occ, clip)
        o.rasterize(occ, clip)
        _same(g, o, report, f"flat facing occluder rasterize<{clip}>")
    V = Views(g, [(128, 64)])
    V.set_mvp(0, mvp)
    V.set_occluders(0, [occ, occ])
    V.run(graph=False)
    V.sync()
    st, rec = V.items(0)
    assert st["c0"].view(np.uint32).tolist() == [0xFFC00000] * 2
    assert st["c1"].view(np.uint32).tolist() == [0xFFC00000] * 2
    valid = (rec["range"] & 0xFFFF) != 0
    assert valid.any() and np.all((rec["modeZ"][valid] >> 16) == 0)
    _, hz = V.raw(0)
    assert np.all(hz[:128] == 1)


def test_vertical_edges_wide_buffer(oracle_mod, report):
    """Exactly vertical edges (edge normal ny == 0) across a 32768x16 buffer,
    where the float32 fold of their offsets over thousands of block columns
    reaches the offset buckets at which the slope rule of Rasterizer.cpp:484
    (ny <= 0) and the table build's (ny < 0) select different rows.  Each
    rectangle is drawn on both sides from a cleared buffer; the sweep's
    rectangles that reach such a bucket on this host (RCPPS sets the edge
    normal's length) are counted and reported."""
    w, h = 32768, 16
    g, o = _pair(oracle_mod, w, h)
    t = o.table().reshape(64, 64)
    differ = {b for b in range(64) if t[62, b] != t[63, b] or t[0, b] != t[1, b]}
    decisive = drawn = 0
    for occ, _, parts in kat.vertical_edge_partials(o, oracle_mod.bake, w, h, range(w - 3000, w - 4, 25)):
        hit = any(int(v) in differ for _, v in parts)
        decisive += hit
        if not hit and drawn % 8:
            drawn += 1
            continue
        drawn += 1
        prod = Occluder(occ.packets, occ.m_refMin, occ.m_refMax, occ.m_boundsMin, occ.m_boundsMax, occ.m_center)
        for r in (g, o):
            r.clear()
        o.rasterize(occ, False)
        g.rasterize(prod, False)
        _same(g, o, report, f"vertical edge at blocks {[bx for bx, _ in parts]}")
    print(f"vertical edges: {drawn} rectangles, {decisive} reach a bucket where the two slope rules differ",
          flush=True)

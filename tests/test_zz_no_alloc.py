"""rz_draw_frame_noalloc: the allocation-free frame draw against rz_draw_frame.

Both entry points are called through the C ABI on two contexts fed the same
calls.  The allocation-free one keeps its working buffers in one static device
array shared by every call on the GPU, so the second draw of each pair runs a
smaller input over what the first left in that array.  Raw depth, hiZ and the
per-occluder statuses must be identical, and the first frame also equals the
oracle's."""

import ctypes

import numpy as np
import pytest

from rasterizer_amd import DeviceBuffer, Rasterizer, lib, scenes

pytestmark = pytest.mark.gpu

RZ_ERR_TOO_LARGE = -6
MAX_ITEMS = 4096


def _noalloc(r: Rasterizer, handles, n: int, d_status: DeviceBuffer) -> None:
    s = lib().rz_draw_frame_noalloc(r.ctx, n, ctypes.cast(handles, ctypes.c_void_p), d_status.ptr)
    assert s == 0, f"rz_draw_frame_noalloc returned {s}"


def _same(a: Rasterizer, b: Rasterizer, report, what: str) -> None:
    da, ha = a.raw()
    db, hb = b.raw()
    assert np.array_equal(ha, hb), f"{what}: hiZ {report(ha, hb)}"
    assert np.array_equal(da, db), f"{what}: depth {report(da, db)}"


def test_noalloc_matches_draw_frame(oracle_mod, report):
    sc = scenes.config1(seed=4)
    m = sc.meta
    w, h = m["width"], m["height"]
    ref = Rasterizer(w, h)
    na = Rasterizer(w, h)
    d_status = DeviceBuffer(na, len(sc.occluders))
    order = scenes.front_to_back(sc.occluders, m["eye"])
    occ = [sc.occluders[i] for i in order]
    # Call 1: the whole scene.  Call 2: a smaller input (every third
    # occluder, another camera) on the buffer call 1 left, after a clear.
    mvp2 = scenes.view_projection([10.0, 5.0, -60.0], [0.1, -0.05, 1.0], [0.0, 1.0, 0.0], np.pi / 3, w / h)
    calls = [("call 1", m["mvp"], occ), ("call 2", mvp2, occ[::3])]
    oracle_st = None
    for what, mvp, items in calls:
        n = len(items)
        print(f"noalloc {what}: {n} occluders, {sum(o.quad_count for o in items)} quads, {w}x{h}", flush=True)
        for r in (ref, na):
            r.clear()
            r.setModelViewProjection(mvp)
        st_ref = ref.draw_frame(items)
        _noalloc(na, na.handle_array(items), n, d_status)
        na.sync()
        st_na = d_status.read(np.uint8, n)
        assert np.array_equal(st_na, st_ref), f"{what}: statuses differ at {np.flatnonzero(st_na != st_ref)[:10]}"
        _same(ref, na, report, what)
        if oracle_st is None:
            o = oracle_mod.OracleRasterizer(w, h)
            o.clear()
            o.setModelViewProjection(mvp)
            oracle_st = o.draw_frame(items)
            assert np.array_equal(st_na, oracle_st), f"{what}: statuses differ from the oracle"
            do, ho = o.raw()
            dn, hn = na.raw()
            assert np.array_equal(hn, ho), f"{what}: hiZ vs oracle {report(hn, ho)}"
            assert np.array_equal(dn, do), f"{what}: depth vs oracle {report(dn, do)}"


def test_noalloc_city_frame(report):
    """The 1920x1080 city frame (285,697 quads in 878 occluders), twice: the
    full frame, then the first 300 occluders of a later frame."""
    sc = scenes.config2(seed=1, frames=16, n_queries=16)
    m = sc.meta
    w, h = m["width"], m["height"]
    ref = Rasterizer(w, h)
    na = Rasterizer(w, h)
    d_status = DeviceBuffer(na, len(sc.occluders))
    for f, count in ((0, len(sc.occluders)), (9, 300)):
        order = scenes.front_to_back(sc.occluders, m["eyes"][f])[:count]
        items = [sc.occluders[i] for i in order]
        print(f"noalloc city frame {f}: {len(items)} occluders", flush=True)
        for r in (ref, na):
            r.clear()
            r.setModelViewProjection(m["mvps"][f])
        st_ref = ref.draw_frame(items)
        _noalloc(na, na.handle_array(items), len(items), d_status)
        na.sync()
        assert np.array_equal(d_status.read(np.uint8, len(items)), st_ref), f"frame {f}: statuses differ"
        _same(ref, na, report, f"city frame {f}")


def test_noalloc_rejects_oversized_input():
    """More occluders than the scratch holds: RZ_ERR_TOO_LARGE, nothing enqueued."""
    sc = scenes.config1(seed=4, n_boxes=64, n_queries=16, width=64, height=64)
    r = Rasterizer(64, 64)
    one = sc.occluders[0]
    handles = r.handle_array([one] * (MAX_ITEMS + 1))
    s = lib().rz_draw_frame_noalloc(r.ctx, MAX_ITEMS + 1, ctypes.cast(handles, ctypes.c_void_p), None)
    assert s == RZ_ERR_TOO_LARGE
    r.sync()

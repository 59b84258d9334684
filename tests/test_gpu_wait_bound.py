"""The bounded cross-tile wait of the query-gated draw, made to give up from
the host.

A query-gated frame makes tiles wait for the other tiles' query2D partials
(Main.cpp:192-206 run on the device).  Each wait is bounded in time; here the
host sets the bound to one tick (10 ns) of the 100 MHz s_memrealtime clock,
so the first wait that finds another tile still drawing gives up.  Every
kernel still exits (the launch ends and the stream drains), the error
surfaces as RZ_ERR_TIMEOUT at the next blocking call, and the context (or
multi-view object) then draws the same frame bit-exactly with the default
bound.  The city frame at 1920x1080 is used because its horizon tiles keep
others waiting for tens of microseconds."""

from __future__ import annotations

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, RasterizerError, Views, scenes

pytestmark = pytest.mark.gpu

RZ_ERR_TIMEOUT = -4


@pytest.fixture(scope="module")
def city():
    return scenes.config2(seed=1, frames=8, n_queries=16)


def _frame(sc, f):
    m = sc.meta
    order = scenes.front_to_back(sc.occluders, m["eyes"][f])
    return m["mvps"][f], [sc.occluders[i] for i in order]


def test_frame_wait_gives_up_and_context_recovers(oracle_mod, city, report):
    m = city.meta
    w, h = m["width"], m["height"]
    mvp, occ = _frame(city, 3)
    g = Rasterizer(w, h)
    nb = (w // 8) * (h // 8)
    g.set_wait_timeout(1)
    g.clear()
    g.setModelViewProjection(mvp)
    print("bounded wait: city frame 3 with a 1-tick bound", flush=True)
    g.draw_frame_handles(g.handle_array(occ), len(occ))
    with pytest.raises(RasterizerError) as e:
        g.sync()
    assert e.value.status == RZ_ERR_TIMEOUT
    g.sync()  # the error word was cleared by the report
    # Default bound again: the same frame from a zeroed buffer equals the oracle.
    g.set_wait_timeout(0)
    g.upload_raw(np.zeros(nb * 64, np.uint16), np.zeros(nb + 8, np.uint16))
    g.clear()
    g.setModelViewProjection(mvp)
    st = g.draw_frame(occ)
    o = oracle_mod.OracleRasterizer(w, h)
    o.clear()
    o.setModelViewProjection(mvp)
    assert np.array_equal(st, o.draw_frame(occ))
    dg, hg = g.raw()
    do, ho = o.raw()
    assert np.array_equal(hg, ho), f"hiZ {report(hg, ho)}"
    assert np.array_equal(dg, do), f"depth {report(dg, do)}"


def test_views_wait_gives_up_and_recovers(oracle_mod, city, report):
    m = city.meta
    w, h = m["width"], m["height"]
    owner = Rasterizer(64, 64)
    frames = (2, 5)
    V = Views(owner, [(w, h)] * len(frames))
    V.set_query_gated(True)
    for k, f in enumerate(frames):
        mvp, occ = _frame(city, f)
        V.set_mvp(k, mvp)
        V.set_occluders(k, occ)
    V.set_wait_timeout(1)
    print("bounded wait: two gated city views with a 1-tick bound", flush=True)
    V.run(graph=False)
    with pytest.raises(RasterizerError) as e:
        V.sync()
    assert e.value.status == RZ_ERR_TIMEOUT
    # Default bound: a step from the buffers the timed-out step left equals
    # the oracle's frame loop from those same buffers.
    V.set_wait_timeout(0)
    before = [V.raw(k) for k in range(len(frames))]
    V.run(graph=False)
    V.sync()
    for k, f in enumerate(frames):
        mvp, occ = _frame(city, f)
        o = oracle_mod.OracleRasterizer(w, h)
        o.set_raw(*before[k])
        o.clear()
        o.setModelViewProjection(mvp)
        st = o.draw_frame(occ)
        assert np.array_equal(V.status(k), st), f"view {k}: statuses"
        dv, hv = V.raw(k)
        do, ho = o.raw()
        assert np.array_equal(hv, ho), f"view {k}: hiZ {report(hv, ho)}"
        assert np.array_equal(dv, do), f"view {k}: depth {report(dv, do)}"

"""The committed golden frame dump (tests/golden/make_frame_golden.py)
reproduced on the device: the single-view path and a one-view query-gated
multi-view launch, from zeroed buffers, bit for bit."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from rasterizer_amd import Rasterizer, Views, load_frame, scenes

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _golden():
    """The fixture for this host's RCPPS: the device tabulates RCPPS from the
    host, as the reference uses it, so a frame is host-specific."""
    sys.path.insert(0, str(GOLDEN))
    import make_frame_golden as g

    path = g.fixture_path()
    if not path.exists():
        pytest.skip(f"no golden frame recorded for this host's RCPPS ({path.name})")
    return g, load_frame(path)


def test_single_view_reproduces_golden_frame():
    g, f = _golden()
    sc = g.scene()
    m = sc.meta
    R = Rasterizer(g.W, g.H)
    R.clear()
    R.setModelViewProjection(m["mvp"])
    R.draw_frame([sc.occluders[i] for i in scenes.front_to_back(sc.occluders, m["eye"])])
    d, h = R.raw()
    v, c = R.query_batch(m["queries"])
    assert np.array_equal(d, f["depth"]) and np.array_equal(h, f["hiz"])
    assert np.array_equal(v, f["visible"]) and np.array_equal(c, f["clip"])


def test_gated_views_reproduce_golden_frame():
    g, f = _golden()
    sc = g.scene()
    m = sc.meta
    V = Views(Rasterizer(64, 64), [(g.W, g.H)])
    V.set_query_gated(True)
    V.set_mvp(0, m["mvp"])
    V.set_occluders(0, list(sc.occluders))
    V.set_eye(0, m["eye"])  # the device sorts front to back
    V.set_queries(0, m["queries"])
    V.run(graph=False)
    V.sync()
    d, h = V.raw(0)
    v, c = V.queries(0)
    assert np.array_equal(d, f["depth"]) and np.array_equal(h, f["hiz"])
    assert np.array_equal(v, f["visible"]) and np.array_equal(c, f["clip"])

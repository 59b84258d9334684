"""On-disk formats (rz_write_occluders / rz_read_occluders, rz_write_frame /
rz_read_frame): round trips, the committed golden frame against the oracle,
and rejection of damaged or foreign files.  Host only."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from rasterizer_amd import RasterizerError, load_frame, load_occluders, save_frame, save_occluders, scenes

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_occluders_round_trip(tmp_path):
    sc = scenes.config1(seed=3, n_boxes=512)
    p = tmp_path / "scene.rzo"
    save_occluders(p, sc.occluders)
    back = load_occluders(p)
    assert len(back) == len(sc.occluders)
    for a, b in zip(sc.occluders, back):
        assert np.array_equal(a.packets, b.packets)
        for f in ("m_refMin", "m_refMax", "m_boundsMin", "m_boundsMax", "m_center"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_empty_occluder_set(tmp_path):
    p = tmp_path / "empty.rzo"
    save_occluders(p, [])
    assert load_occluders(p) == []


def test_frame_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    w, h = 40, 24
    nb = (w // 8) * (h // 8)
    d = rng.integers(0, 65536, nb * 64, dtype=np.uint16)
    z = rng.integers(0, 65536, nb + 8, dtype=np.uint16)
    v = rng.integers(0, 2, 37, dtype=np.uint8)
    c = rng.integers(0, 2, 37, dtype=np.uint8)
    p = tmp_path / "f.rzf"
    save_frame(p, w, h, d, z, v, c)
    f = load_frame(p)
    assert (f["width"], f["height"]) == (w, h)
    assert np.array_equal(f["depth"], d) and np.array_equal(f["hiz"], z)
    assert np.array_equal(f["visible"], v) and np.array_equal(f["clip"], c)
    save_frame(p, w, h, d, z)  # no queries
    f = load_frame(p)
    assert f["visible"].size == 0 and np.array_equal(f["depth"], d)


def test_damaged_and_foreign_files_are_rejected(tmp_path):
    sc = scenes.config1(seed=3, n_boxes=128)
    p = tmp_path / "scene.rzo"
    save_occluders(p, sc.occluders)
    good = p.read_bytes()
    cases = {
        "flipped payload byte": good[:40] + bytes([good[40] ^ 1]) + good[41:],
        "truncated": good[:-9],
        "trailing byte": good + b"\0",
        "other version": good[:8] + (2).to_bytes(4, "little") + good[12:],
        "other magic": b"RZFRM\0\0\0" + good[8:],
    }
    for what, data in cases.items():
        q = tmp_path / "bad.rzo"
        q.write_bytes(data)
        with pytest.raises(RasterizerError) as e:
            load_occluders(q)
        assert e.value.status == -8, what
    with pytest.raises(RasterizerError) as e:
        load_frame(p)  # an occluder file is not a frame dump
    assert e.value.status == -8
    with pytest.raises(RasterizerError) as e:
        load_frame(tmp_path / "missing.rzf")
    assert e.value.status == -8


def _fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for byte in data:
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_frame_header_sizes_are_bounded(tmp_path):
    """A well-formed file (magic, version, size, checksum all valid) whose
    header claims a 2400000008 x 3783947488 buffer and 4110202492 queries:
    12 + blocks*128 + (blocks+8)*2 + queries*2 wraps in 64 bits to exactly the
    76-byte payload.  It must be rejected before any size is trusted."""
    import struct

    w, h, nq = 8 * 300000001, 8 * 472993436, 4110202492
    nb = (w // 8) * (h // 8)
    payload = struct.pack("<III", w, h, nq) + bytes(64)
    assert (12 + nb * 128 + (nb + 8) * 2 + nq * 2) % (1 << 64) == len(payload)
    data = b"RZFRM\0\0\0" + struct.pack("<IIQ", 1, 0, len(payload)) + payload + struct.pack("<Q", _fnv1a64(payload))
    p = tmp_path / "forged.rzf"
    p.write_bytes(data)
    with pytest.raises(RasterizerError) as e:
        load_frame(p)
    assert e.value.status == -8
    # The same header with a real size is accepted (the checksum path is right).
    good = struct.pack("<III", 8, 8, 0) + bytes(128 + 9 * 2)
    p.write_bytes(b"RZFRM\0\0\0" + struct.pack("<IIQ", 1, 0, len(good)) + good + struct.pack("<Q", _fnv1a64(good)))
    assert load_frame(p)["width"] == 8


def test_golden_frame_matches_oracle():
    """The committed frame dump for this host's RCPPS, made by the oracle, is
    what the oracle computes here."""
    sys.path.insert(0, str(GOLDEN))
    import make_frame_golden as g

    path = g.fixture_path()
    if not path.exists():
        pytest.skip(f"no golden frame recorded for this host's RCPPS ({path.name}); "
                    "tests/golden/make_frame_golden.py records one")
    f = load_frame(path)
    d, h, v, c = g.oracle_frame(g.scene())
    assert (f["width"], f["height"]) == (g.W, g.H)
    assert np.array_equal(f["depth"], d) and np.array_equal(f["hiz"], h)
    assert np.array_equal(f["visible"], v) and np.array_equal(f["clip"], c)
    assert int((h[: (g.W // 8) * (g.H // 8)] != 1).sum()) > 100 and 0 < int(v.sum()) < v.size


def test_golden_frames_differ_only_by_host():
    """Every committed golden frame (one per recorded RCPPS) is well formed and
    of the same scene."""
    files = sorted(GOLDEN.glob("frame_config1_s7_128x96_*.rzf"))
    assert files
    for p in files:
        f = load_frame(p)
        assert (f["width"], f["height"]) == (128, 96) and f["visible"].size == 1024

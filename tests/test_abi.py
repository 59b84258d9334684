"""The C ABI library: it loads, exports every entry point that
include/rasterizer_amd.h declares plus the drop-in C++ classes, and its
host-only parts (Occluder::bake, argument checks) behave without a GPU."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import kat

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rasterizer_amd.h"


def _declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(rz_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def rzlib():
    import rasterizer_amd

    return rasterizer_amd.lib()


def test_exports_every_declared_symbol(rzlib):
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(rzlib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    import rasterizer_amd

    names = set(_declared())
    bound = set(rasterizer_amd._SIGNATURES)
    assert names == bound, f"header only: {sorted(names - bound)}, binding only: {sorted(bound - names)}"


def test_dropin_classes_exported():
    """The C++ drop-in (include/rasterizer_amd/Rasterizer.h, Occluder.h) is
    in the same library: every member the reference's headers declare."""
    import rasterizer_amd

    out = subprocess.run(["nm", "-DC", "--defined-only", str(rasterizer_amd.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("Rasterizer::Rasterizer(unsigned int, unsigned int)", "Rasterizer::setModelViewProjection(float const*)",
                "Rasterizer::clear()", "void Rasterizer::rasterize<true>(Occluder const&)",
                "void Rasterizer::rasterize<false>(Occluder const&)", "Rasterizer::queryVisibility(",
                "Rasterizer::query2D(unsigned int, unsigned int, unsigned int, unsigned int, unsigned int) const",
                "Rasterizer::readBackDepth(void*) const", "Occluder::bake(", "Occluder::~Occluder()",
                "rz_dropin::sah_generate_batches("):
        assert sym in out, f"missing drop-in symbol {sym}"


def test_status_strings(rzlib):
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        s = rzlib.rz_status_string(code)
        assert s and len(s) >= 2
    assert rzlib.rz_status_string(12345)


def test_null_arguments_rejected(rzlib):
    """Argument checks run before any device call (no GPU needed)."""
    assert rzlib.rz_set_mvp(None, None) == -1
    assert rzlib.rz_clear(None) == -1
    assert rzlib.rz_rasterize(None, None, 0) == -1
    assert rzlib.rz_bake(None, 0, None, None, None, None, None, None) == -1
    assert rzlib.rz_create(12, 8, None) == -1
    assert rzlib.rz_get_config(None, None) == -1
    assert rzlib.rz_set_config(None, 8, 8) == -1
    assert rzlib.rz_tune_frame(None, 0, None, 1, None, None, None) == -1
    assert rzlib.rz_load_config(None, None, None) == -1
    assert rzlib.rz_views_get_config(None, None) == -1
    assert rzlib.rz_views_set_config(None, 8, 4, 0) == -1
    assert rzlib.rz_views_tune(None, 1, None, None, None) == -1
    assert rzlib.rz_views_load_config(None, None, None) == -1
    assert rzlib.rz_views_read_tile_stats(None, None) == -1
    assert rzlib.rz_sah_batches(None, 0, None, 512, 8, None, None, None) == -1
    assert rzlib.rz_prepare_scene_device(None, 0, None, 512, 8, None, None, None, None, None) == -1
    assert rzlib.rz_occluder_create_device(None, None, 0, None, None, None, None, None) == -1


def _rand_quads(rng, n):
    c = rng.uniform(-20, 20, (n, 1, 3))
    q = c + rng.uniform(-2, 2, (n, 4, 3))
    return np.concatenate([q, np.ones((n, 4, 1))], axis=-1).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bake_matches_oracle(oracle_mod, seed):
    """The library's Occluder::bake (host C++) and the oracle's produce the
    same packets, bounds and centre, including the k-means order."""
    from rasterizer_amd import Occluder

    rng = np.random.default_rng(seed)
    quads = _rand_quads(rng, 8 * (1 + seed * 7))
    v = quads.reshape(-1, 4)
    lo = np.append(v[:, :3].min(axis=0) - 0.5, 1.0).astype(np.float32)
    hi = np.append(v[:, :3].max(axis=0) + 0.5, 1.0).astype(np.float32)
    occ = Occluder.bake(v, lo, hi)
    packets, bmin, bmax, cen = oracle_mod.bake(v, lo, hi)
    assert np.array_equal(occ.packets, packets)
    assert np.array_equal(occ.m_boundsMin, bmin) and np.array_equal(occ.m_boundsMax, bmax)
    assert np.array_equal(occ.m_center, cen)


def test_bake_boxes_match_oracle(oracle_mod):
    """Sixteen boxes (96 faces with six exact axis normals) through both bakes."""
    from rasterizer_amd import Occluder, scenes

    quads = np.concatenate([scenes.box_quads([i, 0, 2 * i], [1, 2, 3]) for i in range(16)], axis=0)
    occ = kat.occluder_from_quads(oracle_mod.bake, quads)
    mine = Occluder.bake(quads.reshape(-1, 4), occ.m_refMin, occ.m_refMax)
    assert np.array_equal(mine.packets, occ.packets)


def test_bake_rejects_partial_batches():
    from rasterizer_amd import Occluder

    with pytest.raises(ValueError):
        Occluder.bake(np.zeros((20, 4), np.float32), np.zeros(4), np.ones(4))

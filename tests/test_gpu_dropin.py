"""The C++ drop-in classes at run time: tests/cpp/main_sequence_check.cpp runs
the reference demo's scene preparation (Main.cpp:86-128: QuadDecomposition,
padding, quad boxes, SurfaceAreaHeuristic::generateBatches, the vertex
buffer's box, Occluder::bake per batch) and its frame (Main.cpp:180-206:
clear, setModelViewProjection, std::sort front to back, queryVisibility then
rasterize<needsClipping> per occluder; readBackDepth, :229) written against
include/rasterizer_amd/*.h exactly as Main.cpp calls them.  Everything it
produces is compared with the reference's own QuadDecomposition and SAH
builds (oracle/_ref) and the oracle's bake and frame loop: occluders bit for
bit, then each frame's readBackDepth bytes, raw depth and hiZ."""

from __future__ import annotations

import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from rasterizer_amd import scenes

pytestmark = pytest.mark.gpu

PROG = Path(__file__).resolve().parent / "_bin" / "main_sequence_check"


def _scene():
    """A 96^2 fBm terrain (2-unit spacing) with 40 boxes on it, one indexed
    triangle mesh as the demo's index and vertex buffers hold scenes."""
    hf = scenes.Heightfield(96, 2.0, 20.0, seed=21)
    tri, verts = scenes.heightfield_triangles(hf)
    rng = np.random.default_rng(5)
    tris, vs, base = [tri], [verts], verts.shape[0]
    for _ in range(40):
        x, z = rng.uniform(10, 180, 2)
        half = rng.uniform(1.0, 6.0, 3)
        c = np.array([x, float(hf.height(x, z)) + half[1] - 0.5, z])
        t, v = scenes.box_mesh(c, half)
        tris.append(t + np.uint32(base))
        vs.append(v)
        base += v.shape[0]
    return hf, np.concatenate(tris).astype(np.uint32), np.concatenate(vs).astype(np.float32)


def _frames(hf, w, h):
    out = []
    for eye, d in (((20.0, 0.0, 30.0), (1.0, -0.08, 0.9)),      # street level: the terrain crosses the near plane
                   ((95.0, 60.0, -40.0), (0.0, -0.6, 1.0)),     # above, looking down the terrain
                   ((170.0, 0.0, 170.0), (-1.0, -0.1, -1.0))):
        e = np.array(eye, np.float64)
        if e[1] == 0.0:
            e[1] = float(hf.height(e[0], e[2])) + 2.0
        d = np.array(d) / np.linalg.norm(d)
        mvp = scenes.view_projection(e, d, (0, 1, 0), math.pi / 3, w / h)
        out.append((mvp, np.array([e[0], e[1], e[2], 0.0], np.float32)))
    return out


def test_main_sequence_through_dropin_headers(oracle_mod, tmp_path, report):
    quad_ref, sah_ref = oracle_mod.quad_reference(), oracle_mod.sah_reference()
    if quad_ref is None or sah_ref is None:
        pytest.skip("the reference's QuadDecomposition / SAH builds (oracle/_ref) are absent")
    assert PROG.exists(), "build the check program with python -m rasterizer_amd.build"
    w, h = 640, 360
    hf, tri, verts = _scene()
    frames = _frames(hf, w, h)
    verts.tofile(tmp_path / "vertices.f32")
    tri.tofile(tmp_path / "indices.u32")
    np.concatenate([np.concatenate([m.reshape(16), c]) for m, c in frames]).astype(np.float32).tofile(
        tmp_path / "frames.f32")
    print(f"main sequence: {tri.shape[0]} triangles, {verts.shape[0]} vertices, {w}x{h}, {len(frames)} frames",
          flush=True)
    run = subprocess.run([str(PROG), str(tmp_path), str(w), str(h), str(len(frames))], capture_output=True,
                         text=True, timeout=300)
    assert run.returncode == 0, run.stderr

    # Expected preparation from the reference's own builds and the oracle.
    idx = quad_ref(tri, verts)
    while idx.size % 32:
        idx = np.append(idx, idx[0])
    q = verts[idx.reshape(-1, 4)]
    batches = sah_ref(q.min(axis=1), q.max(axis=1), 512, 8)
    ref_min, ref_max = verts.min(axis=0), verts.max(axis=0)
    want = []
    for b in batches:
        packets, bmin, bmax, cen = oracle_mod.bake(q[b].reshape(-1, 4), ref_min, ref_max)
        want.append((packets, bmin, bmax, cen))
    raw = np.fromfile(tmp_path / "occluders.bin", np.uint8)
    got, at = [], 0
    while at < raw.size:
        n = int(raw[at:at + 4].view(np.uint32)[0])
        box = raw[at + 4:at + 68].view(np.float32).reshape(4, 4)
        pk = raw[at + 68:at + 68 + 32 * n].view(np.uint32)
        got.append((pk, box))
        at += 68 + 32 * n
    assert len(got) == len(want), f"{len(got)} occluders, expected {len(want)}"
    for i, ((pk, box), (wp, bmin, bmax, _)) in enumerate(zip(got, want)):
        assert np.array_equal(pk, wp), f"occluder {i}: packets"
        assert np.array_equal(box[0].view(np.uint32), bmin.view(np.uint32)), f"occluder {i}: boundsMin"
        assert np.array_equal(box[1].view(np.uint32), bmax.view(np.uint32)), f"occluder {i}: boundsMax"
        assert np.array_equal(box[2], ref_min) and np.array_equal(box[3], ref_max), f"occluder {i}: reference box"

    from types import SimpleNamespace

    occ = [SimpleNamespace(packets=wp, m_refMin=ref_min, m_refMax=ref_max, m_boundsMin=bmin, m_boundsMax=bmax,
                           m_center=cen) for wp, bmin, bmax, cen in want]
    centres = np.stack([o.m_center for o in occ])
    o = oracle_mod.OracleRasterizer(w, h)
    nb = (w // 8) * (h // 8)
    drawn_total = 0
    for k, (mvp, cam) in enumerate(frames):
        data = np.fromfile(tmp_path / f"frame{k}.bin", np.uint8)
        order = data[:4 * len(occ)].view(np.uint32)
        rgba = data[4 * len(occ):4 * len(occ) + w * h * 4]
        rest = data[4 * len(occ) + w * h * 4:].view(np.uint16)
        depth, hiz = rest[:nb * 64], rest[nb * 64:]
        assert sorted(order.tolist()) == list(range(len(occ)))
        # std::sort's order: ascending DPPS-summed squared distance (ties in any order).
        dvec = (centres - cam).astype(np.float32)
        p = (dvec * dvec).astype(np.float32)
        dist = ((p[:, 0] + p[:, 1]).astype(np.float32) + p[:, 2]).astype(np.float32)
        assert np.all(np.diff(dist[order]) >= 0), f"frame {k}: not front to back"
        o.clear()
        o.setModelViewProjection(mvp)
        st = o.draw_frame([occ[i] for i in order])
        drawn_total += int((st > 0).sum())
        do, ho = o.raw()
        assert np.array_equal(hiz, ho), f"frame {k}: hiZ {report(hiz, ho)}"
        assert np.array_equal(depth, do), f"frame {k}: depth {report(depth, do)}"
        assert np.array_equal(rgba, o.readBackDepth().reshape(-1)), f"frame {k}: readBackDepth"
    assert drawn_total > 0

"""Known-answer scene helpers shared by the oracle pins and the GPU parity tests.

Cameras follow the reference demo's conventions (DirectXMath left-handed
look-to and perspective-FOV matrices, Main.cpp:172-178); the screen mapping
of the prebaked matrix (Rasterizer.cpp:76-105) puts NDC x = -1 at prebaked
x' = 0 and NDC x = +1 at x' = width - 8, and pixel column px has its centre at
x' = px - 3.5 (the per-block tables sample (x - 3.5) / 8, Rasterizer.cpp:569-578).
"""

from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from rasterizer_amd import scenes

BIAS = np.float32(2.5237386e-29)  # Rasterizer.cpp:9, 0xFFFF << 12 as a float
NEAR, FAR = 1.0, 5000.0


def camera(width: int, height: int, fov_y: float = math.pi / 3):
    """Eye at the origin looking +z, y up; returns (mvp float[16], proj 4x4 float64)."""
    mvp = scenes.view_projection(np.zeros(3), (0, 0, 1), (0, 1, 0), fov_y, width / height, NEAR, FAR)
    proj = scenes.look_to_lh(np.zeros(3), (0, 0, 1), (0, 1, 0)) @ scenes.perspective_fov_lh(
        fov_y, width / height, NEAR, FAR)
    return mvp, proj


def pack(v: float) -> int:
    """Packed depth of a float32 value: its bits >> 12, saturated to u16."""
    b = int(np.array([v], np.float32).view(np.uint32)[0]) >> 12
    if np.float32(v) < 0:
        return 0
    return min(b, 0xFFFF)


def plane_depth(view_z: float) -> float:
    """Stored depth value of a point at view depth view_z: the prebaked third
    row maps clip (Z, W) to (W - Z) * bias / 2, then divides by W."""
    zc = view_z * FAR / (FAR - NEAR) - NEAR * FAR / (FAR - NEAR)
    return (view_z - zc) * float(BIAS) * 0.5 / view_z


def screen_x(ndc_x: float, width: int) -> float:
    return (ndc_x + 1.0) * (width * 0.5 - 4.0)


def project(proj: np.ndarray, p) -> np.ndarray:
    """Row-vector projection of a world point: returns (ndc_x, ndc_y, view_w)."""
    v = np.array([p[0], p[1], p[2], 1.0]) @ proj
    return np.array([v[0] / v[3], v[1] / v[3], v[3]])


def occluder_from_quads(bake_fn, quads: np.ndarray, ref_min=None, ref_max=None):
    """Bakes quads (n, 4, 4 or 3) padded to whole 8-quad batches by repeating
    the last quad (a repeated quad redraws identical depth, so max-merge keeps
    the result).  bake_fn is oracle.bake or the product's bake."""
    q = np.asarray(quads, np.float32)
    if q.shape[-1] == 3:
        q = np.concatenate([q, np.ones(q.shape[:-1] + (1,), np.float32)], axis=-1)
    rem = (-q.shape[0]) % 8
    if rem:
        q = np.concatenate([q, np.repeat(q[-1:], rem, axis=0)], axis=0)
    v = q.reshape(-1, 4)
    lo = v.min(axis=0) if ref_min is None else np.asarray(ref_min, np.float32)
    hi = v.max(axis=0) if ref_max is None else np.asarray(ref_max, np.float32)
    lo = np.array([lo[0], lo[1], lo[2], 1.0], np.float32)
    hi = np.array([hi[0], hi[1], hi[2], 1.0], np.float32)
    # A flat quad has zero extent on one axis: widen the reference box there.
    for a in range(3):
        if hi[a] <= lo[a]:
            lo[a] -= 1.0
            hi[a] += 1.0
    packets, bmin, bmax, cen = bake_fn(v, lo, hi)
    return SimpleNamespace(packets=packets, m_refMin=lo, m_refMax=hi, m_boundsMin=bmin, m_boundsMax=bmax,
                           m_center=cen)


def facing_rect(x0, x1, y0, y1, z) -> np.ndarray:
    """A quad in the plane view z = const, wound to face a camera at the origin
    looking +z (the -z face of a box, as scenes.box_quads winds it)."""
    c = [(x0 + x1) / 2, (y0 + y1) / 2, z]
    h = [(x1 - x0) / 2, (y1 - y0) / 2, 0.0]
    return scenes.box_quads(c, h)[5]


def unpack_block_linear(depth: np.ndarray, width: int, height: int) -> np.ndarray:
    """Block-linear u16 depth (8x8 blocks, rows of 8) -> (height, width) image."""
    bx, by = width // 8, height // 8
    return depth.reshape(by, bx, 8, 8).transpose(0, 2, 1, 3).reshape(height, width)


def pack_block_linear(img: np.ndarray) -> np.ndarray:
    h, w = img.shape
    return np.ascontiguousarray(img.reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3).reshape(-1))


def vertical_edge_rects(width: int, height: int, pixels, heights=(5.3, 12.7), dist: float = 100.0):
    """Screen-facing rectangles much taller than the screen with one edge on
    screen, exactly vertical in screen space (both of its vertices project to
    the same x'), so that edge's normal has ny == 0 exactly.  For each pixel
    column p in `pixels` and each height (in screen heights), one rectangle
    whose right edge (normal +x) and one whose left edge (normal -x) crosses
    column p; the other vertical edge is far off screen, so the on-screen
    edge's offsets are folded across every block column from the buffer's
    left border.  Returns a list of (quads (1, 4, 4), pixel, edge side)."""
    half_w = dist * math.tan(math.pi / 6) * (width / height)
    half_h = dist * math.tan(math.pi / 6)
    px = 2.0 * half_w / (width - 8.0)  # one x' unit in world units at dist
    out = []
    for p in pixels:
        x = -half_w + (p + 0.37) * px
        for hm in heights:
            y = hm * half_h
            out.append((facing_rect(-4.0 * half_w, x, -y, y, dist), p, "right"))
            out.append((facing_rect(x, 4.0 * half_w, -y, y, dist), p, "left"))
    return out


def vertical_edge_partials(o, bake_fn, width: int, height: int, pixels):
    """For each rectangle of vertical_edge_rects: its on-screen vertical
    edge's per-block offsets, folded from the rectangle's first block column
    in float32 as the traversal folds them (Rasterizer.cpp:1136-1146: one add
    of the edge normal's x per block), from the setup of the oracle
    rasterizer `o` (which gets the camera's matrix), and the blocks where
    they are partial (0 <= offset < 63).  Yields (occluder, slope row 2q,
    [(block x, offset)])."""
    mvp, _ = camera(width, height)
    o.setModelViewProjection(mvp)
    for quad, _, _ in vertical_edge_rects(width, height, pixels):
        occ = occluder_from_quads(bake_fn, [quad])
        d = o.setup_dump(occ, False)[0]
        assert d[0] == 1 and d[1] == 6, "a screen-facing rectangle is a valid Convex quad"
        min_x, range_x = int(d[2]), int(d[4])
        off, nx, ny = d[8:12].astype(np.float32), d[12:16].astype(np.float32), d[16:20].astype(np.float32)
        for e in range(4):
            if ny[e] != 0 or range_x < 1:
                continue
            seq = np.add.accumulate(np.concatenate([[off[e]], np.full(range_x - 1, nx[e], np.float32)]),
                                    dtype=np.float32)
            part = np.flatnonzero((seq >= 0) & (seq < 63))
            # nx * mul + add is 31.5 (+x) or 0.5 (-x) before truncation.
            yield occ, (62 if nx[e] > 0 else 0), [(min_x + int(j), float(seq[j])) for j in part]

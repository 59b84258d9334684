"""Seeded synthetic scenes and cameras for the rasterizer's tests and benchmark.

Everything here is host-side scene preparation, restating what the
reference's demo does before its frame loop (Main.cpp:86-128):
quads -> pad to whole 8-quad batches -> SAH batches of < 512 quads
(SurfaceAreaHeuristic.cpp:10-104) -> Occluder::bake with the scene AABB.
Random numbers come from SplitMix64 so scenes are identical on every host.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import Occluder

# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); floats from the top 24 bits."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, n: int) -> np.ndarray:
        """n floats in [0, 1), float32, = (u >> 40) * 2^-24."""
        out = np.empty(n, dtype=np.float64)
        for i in range(n):
            out[i] = (self.next_u64() >> 40) * (1.0 / 16777216.0)
        return out.astype(np.float32)

    def uniform_range(self, n: int, lo: float, hi: float) -> np.ndarray:
        u = self.uniform(n).astype(np.float64)
        return (lo + (hi - lo) * u).astype(np.float32)

    def log_uniform(self, n: int, lo: float, hi: float) -> np.ndarray:
        u = self.uniform(n).astype(np.float64)
        return np.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u).astype(np.float32)


# ---------------------------------------------------------------------------
# Camera (DirectXMath's documented LH matrices, row-vector convention)
# ---------------------------------------------------------------------------
def look_to_lh(eye, direction, up) -> np.ndarray:
    """XMMatrixLookToLH: rows (R0.x R1.x R2.x 0), ..., (-R0.eye -R1.eye -R2.eye 1)."""
    eye = np.asarray(eye, np.float64)
    r2 = np.asarray(direction, np.float64)
    r2 = r2 / np.linalg.norm(r2)
    r0 = np.cross(np.asarray(up, np.float64), r2)
    r0 = r0 / np.linalg.norm(r0)
    r1 = np.cross(r2, r0)
    m = np.zeros((4, 4))
    m[:3, 0], m[:3, 1], m[:3, 2] = r0, r1, r2
    m[3, :3] = [-r0 @ eye, -r1 @ eye, -r2 @ eye]
    m[3, 3] = 1.0
    return m


def perspective_fov_lh(fov_y: float, aspect: float, near: float, far: float) -> np.ndarray:
    """XMMatrixPerspectiveFovLH: depth 0 at near, 1 at far."""
    h = math.cos(0.5 * fov_y) / math.sin(0.5 * fov_y)
    w = h / aspect
    rng = far / (far - near)
    m = np.zeros((4, 4))
    m[0, 0], m[1, 1], m[2, 2], m[2, 3], m[3, 2] = w, h, rng, 1.0, -rng * near
    return m


def view_projection(eye, direction, up, fov_y, aspect, near=1.0, far=5000.0) -> np.ndarray:
    """float[16] as Main.cpp:172-178 memcpy's the XMMATRIX (row-major)."""
    return (look_to_lh(eye, direction, up) @ perspective_fov_lh(fov_y, aspect, near, far)).astype(
        np.float32).reshape(16)


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------
def box_quads(center, half) -> np.ndarray:
    """Six outward faces of an axis-aligned box as quads (6, 4, 4), w = 1.

    Each face is wound so that, seen from outside, its screen-space area is
    positive: (v1 - v0) x (v2 - v1) points into the box (the reference culls
    area <= 0 as back-facing, Rasterizer.cpp:775-791)."""
    c = np.asarray(center, np.float64)
    h = np.asarray(half, np.float64)
    out = np.zeros((6, 4, 4), np.float32)
    f = 0
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        for sign in (1.0, -1.0):
            n = np.zeros(3)
            n[axis] = sign
            corners = []
            for su, sv in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                p = c.copy()
                p[axis] += sign * h[axis]
                p[u] += su * h[u]
                p[v] += sv * h[v]
                corners.append(p)
            cr = np.cross(corners[1] - corners[0], corners[2] - corners[1])
            if cr @ n > 0:
                corners = corners[::-1]
            for k in range(4):
                out[f, k, :3] = corners[k]
                out[f, k, 3] = 1.0
            f += 1
    return out


def pad_to_batches(quads: np.ndarray) -> np.ndarray:
    """Pads to a multiple of 8 quads with degenerate copies of the first
    vertex, as Main.cpp:90-94 pads its index buffer with indices[0]."""
    n = quads.shape[0]
    rem = (-n) % 8
    if rem == 0:
        return quads
    pad = np.repeat(quads[:1, :1, :], 4, axis=1)
    return np.concatenate([quads, np.repeat(pad, rem, axis=0)], axis=0)


def _surface_area(mn: np.ndarray, mx: np.ndarray) -> np.ndarray:
    """Aabb::surfaceArea (VectorMath.h:57-62): DPPS(ext, ext.yzx, 0x7F)."""
    e = (mx - mn).astype(np.float32)
    p0 = (e[..., 0] * e[..., 1]).astype(np.float32)
    p1 = (e[..., 1] * e[..., 2]).astype(np.float32)
    p2 = (e[..., 2] * e[..., 0]).astype(np.float32)
    return ((p0 + p1).astype(np.float32) + p2).astype(np.float32)


def _sah_split(cmin, cmax, centers, idx, gran):
    n = idx.size
    best_cost, best_axis, best_index = np.float32(np.inf), -1, -1
    for axis in range(3):
        idx = idx[np.argsort(centers[idx, axis], kind="stable")]
        lmin = np.minimum.accumulate(cmin[idx], axis=0)
        lmax = np.maximum.accumulate(cmax[idx], axis=0)
        rmin = np.minimum.accumulate(cmin[idx][::-1], axis=0)[::-1]
        rmax = np.maximum.accumulate(cmax[idx][::-1], axis=0)[::-1]
        left = _surface_area(lmin, lmax)
        right = _surface_area(rmin, rmax)
        splits = np.arange(gran, n - gran, gran)
        if splits.size == 0:
            continue
        cost = (left[splits - 1] * splits.astype(np.float32)).astype(np.float32) + \
               (right[splits] * (n - splits).astype(np.float32)).astype(np.float32)
        k = int(np.argmin(cost))
        if cost[k] < best_cost:
            best_cost, best_axis, best_index = cost[k], axis, int(splits[k])
    if best_axis < 0:
        return idx, n
    idx = idx[np.argsort(centers[idx, best_axis], kind="stable")]
    return idx, best_index


def sah_batches(quad_min: np.ndarray, quad_max: np.ndarray, target: int = 512, gran: int = 8) -> list:
    """SurfaceAreaHeuristic::generateBatches (SurfaceAreaHeuristic.cpp:75-104):
    recursive SAH splits at multiples of `gran` until every batch holds fewer
    than `target` quads.  Returns lists of quad indices in batch order."""
    cmin = np.asarray(quad_min, np.float32)[:, :3]
    cmax = np.asarray(quad_max, np.float32)[:, :3]
    centers = (cmin + cmax).astype(np.float32)
    out = []

    def rec(idx):
        if idx.size < target:
            out.append(idx)
            return
        idx, split = _sah_split(cmin, cmax, centers, idx, gran)
        for part in (idx[:split], idx[split:]):
            if part.size == 0:
                continue
            if part.size < target:
                out.append(part)
            else:
                rec(part)

    rec(np.arange(cmin.shape[0]))
    return out


# ---------------------------------------------------------------------------
# Scenes
# ---------------------------------------------------------------------------
@dataclass
class Scene:
    occluders: list
    quad_count: int          # real quads (padding excluded)
    ref_min: np.ndarray
    ref_max: np.ndarray
    meta: dict = field(default_factory=dict)


def bake_scene(quads: np.ndarray, real_quads: int | None = None, target: int = 512, gran: int = 8) -> Scene:
    """Main.cpp:90-128: pad, SAH-batch and bake with the scene AABB."""
    q = pad_to_batches(np.asarray(quads, np.float32))
    verts = q.reshape(-1, 4)
    ref_min = verts.min(axis=0).astype(np.float32)
    ref_max = verts.max(axis=0).astype(np.float32)
    batches = sah_batches(q.min(axis=1), q.max(axis=1), target, gran)
    occ = []
    for b in batches:
        occ.append(Occluder.bake(q[b].reshape(-1, 4), ref_min, ref_max))
    return Scene(occ, int(real_quads if real_quads is not None else quads.shape[0]), ref_min, ref_max)


def random_boxes(rng: SplitMix64, n: int, lo: float, hi: float, hmin: float, hmax: float):
    centers = np.stack([rng.uniform_range(n, lo, hi) for _ in range(3)], axis=1)
    half = np.stack([rng.log_uniform(n, hmin, hmax) for _ in range(3)], axis=1)
    return centers, half


def aabb_bounds(centers: np.ndarray, half: np.ndarray) -> np.ndarray:
    """Query boxes as (n, 8) float32: min xyz 1, max xyz 1."""
    n = centers.shape[0]
    b = np.ones((n, 8), np.float32)
    b[:, 0:3] = centers - half
    b[:, 4:7] = centers + half
    return b


def config1(seed: int = 1, n_boxes: int = 1024, n_queries: int = 4096):
    """1,024 box occluders (centres U[-50,50]^3, half-extents log-U[0.5,5]),
    camera at the origin looking +z (60 deg vertical FOV, near 1, far 5000,
    1280x720), then 4,096 query boxes (half-extents log-U[0.1,2])."""
    rng = SplitMix64(seed)
    c, h = random_boxes(rng, n_boxes, -50.0, 50.0, 0.5, 5.0)
    quads = np.concatenate([box_quads(c[i], h[i]) for i in range(n_boxes)], axis=0)
    scene = bake_scene(quads)
    qc, qh = random_boxes(rng, n_queries, -50.0, 50.0, 0.1, 2.0)
    width, height = 1280, 720
    eye = np.zeros(3, np.float32)
    mvp = view_projection(eye, (0, 0, 1), (0, 1, 0), math.pi / 3, width / height)
    scene.meta.update(width=width, height=height, eye=eye, mvp=mvp, queries=aabb_bounds(qc, qh), seed=seed)
    return scene


def front_to_back(occluders, eye) -> np.ndarray:
    """Draw order of Main.cpp:184-190: ascending squared distance from the eye
    to each occluder's centre (float32, DPPS order), ties by index."""
    cen = np.stack([np.asarray(o.m_center, np.float32)[:3] for o in occluders])
    d = (cen - np.asarray(eye, np.float32)[:3]).astype(np.float32)
    p = (d * d).astype(np.float32)
    dist = ((p[:, 0] + p[:, 1]).astype(np.float32) + p[:, 2]).astype(np.float32)
    return np.argsort(dist, kind="stable")

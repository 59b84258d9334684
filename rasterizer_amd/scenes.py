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
    assert best_axis >= 0, "a part that is split has more than 2 * gran quads"
    idx = idx[np.argsort(centers[idx, best_axis], kind="stable")]
    return idx, best_index


def sah_batches(quad_min: np.ndarray, quad_max: np.ndarray, target: int = 512, gran: int = 8) -> list:
    """SurfaceAreaHeuristic::generateBatches (SurfaceAreaHeuristic.cpp:75-104):
    recursive SAH splits at multiples of `gran`.  The whole set is always
    split (generateBatches calls the recursion on it directly); a part with
    fewer than `target` quads is a batch, a larger one is split again.
    Returns lists of quad indices in batch order.  Needs target > 2 * gran
    (so that every part that is split again has a split position); a set of
    at most 2 * gran quads, where the reference's split is undefined, is one
    batch in input order."""
    if gran < 1 or target <= 2 * gran:
        raise ValueError("sah_batches needs gran >= 1 and target > 2 * gran")
    cmin = np.asarray(quad_min, np.float32)[:, :3]
    cmax = np.asarray(quad_max, np.float32)[:, :3]
    n = cmin.shape[0]
    if n == 0:
        return []
    if n <= 2 * gran:
        return [np.arange(n)]
    centers = (cmin + cmax).astype(np.float32)
    out = []

    def rec(idx):
        idx, split = _sah_split(cmin, cmax, centers, idx, gran)
        for part in (idx[:split], idx[split:]):
            if part.size < target:
                out.append(part)
            else:
                rec(part)

    rec(np.arange(n))
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


def bake_scene(quads: np.ndarray, real_quads: int | None = None, target: int = 512, gran: int = 8,
               vertex_buffer: np.ndarray | None = None) -> Scene:
    """Main.cpp:90-128: pad, SAH-batch and bake with the scene AABB, taken
    over vertex_buffer (Main.cpp:109-113 takes it over the whole vertex
    buffer) or, without one, over the quads' vertices."""
    q = pad_to_batches(np.asarray(quads, np.float32))
    verts = q.reshape(-1, 4) if vertex_buffer is None else np.asarray(vertex_buffer, np.float32).reshape(-1, 4)
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


def config1(seed: int = 1, n_boxes: int = 1024, n_queries: int = 4096, width: int = 1280, height: int = 720):
    """1,024 box occluders (centres U[-50,50]^3, half-extents log-U[0.5,5]),
    camera at the origin looking +z (60 deg vertical FOV, near 1, far 5000,
    1280x720), then 4,096 query boxes (half-extents log-U[0.1,2])."""
    rng = SplitMix64(seed)
    c, h = random_boxes(rng, n_boxes, -50.0, 50.0, 0.5, 5.0)
    quads = np.concatenate([box_quads(c[i], h[i]) for i in range(n_boxes)], axis=0)
    scene = bake_scene(quads)
    qc, qh = random_boxes(rng, n_queries, -50.0, 50.0, 0.1, 2.0)
    eye = np.zeros(3, np.float32)
    mvp = view_projection(eye, (0, 0, 1), (0, 1, 0), math.pi / 3, width / height)
    scene.meta.update(width=width, height=height, eye=eye, mvp=mvp, queries=aabb_bounds(qc, qh), seed=seed)
    return scene


def _mix64(h: np.ndarray) -> np.ndarray:
    """SplitMix64 finalizer on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        h = (h ^ (h >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        h = (h ^ (h >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return h ^ (h >> np.uint64(31))


def _lattice(ix: np.ndarray, iz: np.ndarray, seed: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        h = (ix.astype(np.int64).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ \
            (iz.astype(np.int64).astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ np.uint64(seed & _M64)
    return (_mix64(h) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def fbm(x: np.ndarray, z: np.ndarray, seed: int, octaves: int = 4, base_period: float = 256.0) -> np.ndarray:
    """Fractal sum of smoothstep-interpolated value noise in [-1, 1]."""
    total = np.zeros(np.broadcast(x, z).shape)
    amp, norm, freq = 1.0, 0.0, 1.0 / base_period
    for o in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx), np.floor(fz)
        tx, tz = fx - ix, fz - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        s = seed + 7919 * o
        v00, v10 = _lattice(ix, iz, s), _lattice(ix + 1, iz, s)
        v01, v11 = _lattice(ix, iz + 1, s), _lattice(ix + 1, iz + 1, s)
        v = (v00 * (1 - sx) + v10 * sx) * (1 - sz) + (v01 * (1 - sx) + v11 * sx) * sz
        total += amp * (2.0 * v - 1.0)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


class Heightfield:
    """n x n vertices at `spacing`, y = amplitude * fbm(x, z)."""

    def __init__(self, n: int, spacing: float, amplitude: float, seed: int):
        self.n, self.spacing = n, spacing
        g = np.arange(n) * spacing
        X, Z = np.meshgrid(g, g, indexing="ij")
        self.h = (amplitude * fbm(X, Z, seed)).astype(np.float32)  # h[i, j] at x = i*s, z = j*s

    def height(self, x, z):
        """Bilinear height at world (x, z), clamped to the grid."""
        s, n = self.spacing, self.n
        fx = np.clip(np.asarray(x, np.float64) / s, 0, n - 1 - 1e-9)
        fz = np.clip(np.asarray(z, np.float64) / s, 0, n - 1 - 1e-9)
        i, j = np.floor(fx).astype(int), np.floor(fz).astype(int)
        tx, tz = fx - i, fz - j
        h = self.h.astype(np.float64)
        return (h[i, j] * (1 - tx) + h[i + 1, j] * tx) * (1 - tz) + (h[i, j + 1] * (1 - tx) + h[i + 1, j + 1] * tx) * tz

    def quads(self) -> np.ndarray:
        """The triangulated grid (heightfield_triangles) through the
        reference's quad decomposition (QuadDecomposition.cpp:346-445, the
        library's rz_quad_decompose): quads (m, 4, 4), up-facing surfaces
        with positive screen-space area."""
        tri, verts = heightfield_triangles(self)
        return decompose_mesh(tri, verts)


def heightfield_triangles(hf: Heightfield) -> tuple[np.ndarray, np.ndarray]:
    """Indexed triangle mesh of a heightfield: vertex i * n + j at
    (i * spacing, h[i, j], j * spacing, 1); each cell split along its
    (i, j)-(i+1, j+1) diagonal into two triangles, cell by cell.  Triangles
    are wound as the reference's index buffers are: opposite to the quads
    QuadDecomposition emits, which write a pair (i0, i1, i2) + apex as
    (i0, i2, i1, apex) (QuadDecomposition.cpp:420-427)."""
    n, s = hf.n, hf.spacing
    g = np.arange(n, dtype=np.float64) * s
    X, Z = np.meshgrid(g, g, indexing="ij")
    verts = np.stack([X.ravel(), hf.h.astype(np.float64).ravel(), Z.ravel(), np.ones(n * n)], axis=1)
    ij = np.arange(n * n, dtype=np.uint32).reshape(n, n)
    p00, p10, p11, p01 = ij[:-1, :-1].ravel(), ij[1:, :-1].ravel(), ij[1:, 1:].ravel(), ij[:-1, 1:].ravel()
    tri = np.stack([np.stack([p00, p11, p10], 1), np.stack([p00, p01, p11], 1)], 1).reshape(-1, 3)
    return tri.astype(np.uint32), verts.astype(np.float32)


def box_mesh(center, half) -> tuple[np.ndarray, np.ndarray]:
    """An axis-aligned box as an indexed mesh: 8 vertices, 12 triangles wound
    so that QuadDecomposition pairs each face's two triangles back into
    box_quads' outward-facing quads."""
    q = box_quads(center, half)
    verts, index = [], {}
    tri = []
    for f in range(6):
        ids = []
        for k in range(4):
            key = tuple(q[f, k, :3].tolist())
            if key not in index:
                index[key] = len(verts)
                verts.append(q[f, k])
            ids.append(index[key])
        a, b, c, d = ids
        tri += [(a, c, b), (a, d, c)]
    return np.array(tri, np.uint32), np.array(verts, np.float32)


def triangulated_boxes(rng: np.random.Generator, n: int) -> tuple[np.ndarray, np.ndarray]:
    """n random boxes (centres U[-50,50]^3, half-extents U[1,5]) as one
    indexed triangle mesh."""
    tris, verts, base = [], [], 0
    for _ in range(n):
        t, v = box_mesh(rng.uniform(-50, 50, 3), rng.uniform(1, 5, 3))
        tris.append(t + base)
        verts.append(v)
        base += v.shape[0]
    return np.concatenate(tris).astype(np.uint32), np.concatenate(verts).astype(np.float32)


def QuadDecomposition_indices(tri: np.ndarray, verts: np.ndarray) -> np.ndarray:
    """QuadDecomposition::decompose (the library's rz_quad_decompose): the
    quads' vertex indices, (m, 4)."""
    from . import QuadDecomposition

    return QuadDecomposition.decompose(tri, verts).reshape(-1, 4).astype(np.int64)


def decompose_mesh(tri: np.ndarray, verts: np.ndarray) -> np.ndarray:
    """QuadDecomposition::decompose of an indexed triangle mesh, as quads
    (m, 4, 4) of the mesh's vertices."""
    return np.ascontiguousarray(np.asarray(verts, np.float32)[QuadDecomposition_indices(tri, verts)])


def catmull_rom(points: np.ndarray, samples: int):
    """Uniform Catmull-Rom through points[1..-2]; returns positions and tangents."""
    P = np.asarray(points, np.float64)
    segs = P.shape[0] - 3
    t = np.linspace(0.0, segs, samples, endpoint=False)
    k = np.minimum(t.astype(int), segs - 1)
    u = (t - k)[:, None]
    p0, p1, p2, p3 = P[k], P[k + 1], P[k + 2], P[k + 3]
    pos = 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u ** 2 + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)
    tan = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * u + 3 * (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 2)
    return pos, tan


def config2(seed: int = 1, frames: int = 256, n_queries: int = 65536, grid: int = 64, terrain_n: int = 512,
            width: int = 1920, height: int = 1080):
    """Procedural city on terrain: grid x grid box buildings (footprint 8x8 on
    a 16-unit pitch, heights log-U[5,80]) on a terrain_n^2 fBm heightfield
    (amplitude 20, 4 octaves, 2-unit spacing), triangulated as one indexed
    mesh and quad-decomposed (QuadDecomposition::decompose), ~286k quads at
    the defaults, baked against the vertex buffer's box (Main.cpp:86-128); a
    seeded Catmull-Rom path along the streets at eye height 2 (near plane 1,
    so terrain quads cross it); query boxes are the building bounds plus
    small props on the terrain."""
    rng = SplitMix64(seed)
    spacing = 2.0
    hf = Heightfield(terrain_n, spacing, 20.0, seed=(seed * 0x2545F4914F6CDD1D) & _M64)
    extent = (terrain_n - 1) * spacing
    pitch = extent / grid
    heights = rng.log_uniform(grid * grid, 5.0, 80.0)
    terrain_tri, terrain_v = heightfield_triangles(hf)
    tris, verts, bb = [terrain_tri], [terrain_v], []
    vbase = terrain_v.shape[0]
    for i in range(grid):
        for j in range(grid):
            cx, cz = (i + 0.5) * pitch, (j + 0.5) * pitch
            xs = np.array([cx - 4, cx + 4, cx - 4, cx + 4, cx])
            zs = np.array([cz - 4, cz - 4, cz + 4, cz + 4, cz])
            base = float(hf.height(xs, zs).min()) - 1.0
            hgt = float(heights[i * grid + j])
            c = np.array([cx, base + 0.5 * hgt, cz])
            h = np.array([4.0, 0.5 * hgt, 4.0])
            t, v = box_mesh(c, h)
            tris.append(t + np.uint32(vbase))
            verts.append(v)
            vbase += v.shape[0]
            bb.append((c, h))
    # The whole scene as one indexed triangle mesh, quad-decomposed and baked
    # as Main.cpp:86-128 prepares its scene.
    tri = np.concatenate(tris).astype(np.uint32)
    vbuf = np.concatenate(verts).astype(np.float32)
    qidx = QuadDecomposition_indices(tri, vbuf)
    quads = vbuf[qidx]
    terrain_quads = int((qidx < terrain_v.shape[0]).all(axis=1).sum())
    scene = bake_scene(quads, vertex_buffer=vbuf)

    # Query boxes: building bounds, then props on the terrain.
    nb = len(bb)
    n_props = max(n_queries - nb, 0)
    px = rng.uniform_range(n_props, 0.0, extent)
    pz = rng.uniform_range(n_props, 0.0, extent)
    ph = np.stack([rng.log_uniform(n_props, 0.25, 1.5) for _ in range(3)], axis=1)
    py = (hf.height(px, pz) + ph[:, 1]).astype(np.float32)
    centers = np.concatenate([np.array([c for c, _ in bb], np.float32), np.stack([px, py, pz], axis=1)])
    halves = np.concatenate([np.array([h for _, h in bb], np.float32), ph])
    queries = aabb_bounds(centers, halves)[:n_queries]

    # Camera path: a walk over street intersections, smoothed.
    pts = []
    ci, cj = 1 + int(rng.uniform(1)[0] * (grid - 2)), 1 + int(rng.uniform(1)[0] * (grid - 2))
    for step in range(12):
        pts.append((ci * pitch, cj * pitch))
        d = 2 + int(rng.uniform(1)[0] * 5)
        sgn = 1 if rng.uniform(1)[0] < 0.5 else -1
        if step % 2 == 0:
            ci = int(np.clip(ci + sgn * d, 1, grid - 1))
        else:
            cj = int(np.clip(cj + sgn * d, 1, grid - 1))
    pos2, tan2 = catmull_rom(np.array(pts), frames)
    eyes = np.zeros((frames, 3), np.float32)
    eyes[:, 0], eyes[:, 2] = pos2[:, 0], pos2[:, 1]
    eyes[:, 1] = hf.height(pos2[:, 0], pos2[:, 1]) + 2.0
    mvps = np.zeros((frames, 16), np.float32)
    for f in range(frames):
        d = np.array([tan2[f, 0], 0.0, tan2[f, 1]])
        nrm = np.linalg.norm(d)
        d = d / nrm if nrm > 1e-9 else np.array([1.0, 0.0, 0.0])
        d[1] = -0.08
        mvps[f] = view_projection(eyes[f], d, (0, 1, 0), math.pi / 3, width / height)
    scene.meta.update(width=width, height=height, eyes=eyes, mvps=mvps, queries=queries, seed=seed,
                      heightfield=hf, pitch=pitch, grid=grid,
                      terrain_quads=terrain_quads, building_quads=int(quads.shape[0] - terrain_quads))
    return scene


CONFIG3_SIZES = [(2048, 2048)] * 16 + [(1024, 1024)] * 24 + [(512, 512)] * 16 + [(1920, 1080)] * 8


def config3(seed: int = 3, n_boxes: int = 40000, half_cube: float = 171.0, n_queries: int = 16384,
            sizes=None):
    """Batched multi-view: config 1's box distribution (half-extents
    log-U[0.5,5]) over a cube of the same box density but 40,000 boxes
    (centres U[-171,171]^3, 240,000 quads), so that per-view subsets of up to
    200k quads exist.  Views (default 64: 16 at 2048^2, 24 at 1024^2, 16 at
    512^2, 8 at 1920x1080) each get an eye inside the cube, a direction, an
    occluder subset whose real quad count is log-U[1k,200k] in front-to-back
    order for that eye, and 16,384 query boxes of config 1's query
    distribution around the eye."""
    rng = SplitMix64(seed)
    c, h = random_boxes(rng, n_boxes, -half_cube, half_cube, 0.5, 5.0)
    quads = np.concatenate([box_quads(c[i], h[i]) for i in range(n_boxes)], axis=0)
    scene = bake_scene(quads)
    sizes = list(CONFIG3_SIZES if sizes is None else sizes)
    pick = np.random.default_rng(seed)
    real = np.array([o.quad_count for o in scene.occluders])
    views = []
    for w, hh in sizes:
        eye = pick.uniform(-0.8 * half_cube, 0.8 * half_cube, 3).astype(np.float32)
        d = pick.normal(size=3)
        d /= np.linalg.norm(d)
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.95 else (1.0, 0.0, 0.0)
        budget = math.exp(pick.uniform(math.log(1000.0), math.log(200000.0)))
        order = pick.permutation(len(scene.occluders))
        k = int(np.searchsorted(np.cumsum(real[order]), budget)) + 1
        sub = [scene.occluders[i] for i in order[:k]]
        sub = [sub[i] for i in front_to_back(sub, eye)]
        qc = eye + pick.uniform(-50.0, 50.0, (n_queries, 3))
        qh = np.exp(pick.uniform(math.log(0.1), math.log(2.0), (n_queries, 3)))
        views.append(dict(size=(w, hh), eye=eye, dir=d.astype(np.float32), up=up, occluders=sub,
                          quads=int(sum(o.quad_count for o in sub)), queries=aabb_bounds(qc, qh)))
    scene.meta.update(views=views, seed=seed)
    return scene


def view_mvp(view: dict, step: int = 0, speed: float = 0.5) -> np.ndarray:
    """Matrix of a config-3 view at a step: the eye moves along its direction."""
    w, h = view["size"]
    eye = view["eye"] + np.float32(speed * step) * view["dir"]
    return view_projection(eye, view["dir"], view["up"], math.pi / 3, w / h)


def front_to_back(occluders, eye) -> np.ndarray:
    """Draw order of Main.cpp:184-190: ascending squared distance from the eye
    to each occluder's centre (float32, DPPS order), ties by index."""
    cen = np.stack([np.asarray(o.m_center, np.float32)[:3] for o in occluders])
    d = (cen - np.asarray(eye, np.float32)[:3]).astype(np.float32)
    p = (d * d).astype(np.float32)
    dist = ((p[:, 0] + p[:, 1]).astype(np.float32) + p[:, 2]).astype(np.float32)
    return np.argsort(dist, kind="stable")


def icosphere() -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere with 80 triangles: the icosahedron with every face split
    into four, new vertices pushed to the sphere.  Returns (vertices (42, 3),
    faces (80, 3)), faces wound as box_quads' faces: front-facing seen from
    outside under the left-handed view matrices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    mid = {}

    def midpoint(a, b):
        k = (min(a, b), max(a, b))
        if k not in mid:
            m = verts[a] + verts[b]
            verts.append(m / np.linalg.norm(m))
            mid[k] = len(verts) - 1
        return mid[k]

    faces = []
    for a, b, c in f:
        ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
        faces += [(a, ca, ab), (b, ab, bc), (c, bc, ca), (ab, ca, bc)]
    return np.array(verts), np.array(faces)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """n uniformly random rotation matrices (Shoemake's unit quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


CONFIG5_SIZES = [(3840, 2160)] + [(1024, 1024)] * 15


def config5(seed: int = 5, n_instances: int = 50000, terrain_n: int = 512, extent: float = 2000.0,
            queries_per_view: int = 65536, sizes=None):
    """Stress scene: n_instances seeded instances of the 80-triangle icosphere
    (scales log-U[0.5,10], positions U over an extent x extent square resting
    on the terrain, uniformly random rotations; every triangle one degenerate
    quad (a, b, c, a), as leftover triangles of the reference's quad
    decomposition are) on config 2's fBm terrain stretched to the same square.
    Each instance is one occluder baked against its own bounding box (80
    quads, 10 batches); the terrain is SAH-batched against its bounds.
    Views (default 16): one street-level 3840x2160 main view and 15
    shadow-like 1024^2 views looking down from 300-700 units above the
    terrain; each gets queries_per_view query boxes (instance bounds near its
    target plus props on the terrain)."""
    sizes = list(CONFIG5_SIZES if sizes is None else sizes)
    rng = np.random.default_rng(seed)
    spacing = extent / (terrain_n - 1)
    hf = Heightfield(terrain_n, spacing, 20.0, seed=(seed * 0x2545F4914F6CDD1D) & _M64)
    terrain = hf.quads()
    tq = pad_to_batches(terrain)
    tv = tq.reshape(-1, 4)
    t_min, t_max = tv.min(axis=0).astype(np.float32), tv.max(axis=0).astype(np.float32)
    occ = [Occluder.bake(tq[b].reshape(-1, 4), t_min, t_max)
           for b in sah_batches(tq.min(axis=1), tq.max(axis=1), 512, 8)]
    n_terrain = len(occ)

    base_v, base_f = icosphere()
    scale = np.exp(rng.uniform(math.log(0.5), math.log(10.0), n_instances))
    px = rng.uniform(0.0, extent, n_instances)
    pz = rng.uniform(0.0, extent, n_instances)
    py = hf.height(px, pz) + scale
    centers = np.stack([px, py, pz], axis=1)
    R = _random_rotations(rng, n_instances)
    tri = base_v[base_f]                                   # (80, 3, 3)
    for k in range(n_instances):
        p = (tri @ R[k].T) * scale[k] + centers[k]         # (80, 3, 3)
        q = np.ones((80, 4, 4), np.float32)
        q[:, 0:3, 0:3] = p
        q[:, 3, 0:3] = p[:, 0]                             # (a, b, c, a)
        v = q.reshape(-1, 4)
        occ.append(Occluder.bake(v, v.min(axis=0), v.max(axis=0)))
    scene = Scene(occ, int(terrain.shape[0] + 80 * n_instances), t_min, t_max)

    views = []
    for i, (w, h) in enumerate(sizes):
        if i == 0 and w >= 2048:
            eye = np.array([0.5 * extent, 0.0, 0.05 * extent])
            eye[1] = hf.height(eye[0], eye[2]) + 2.0
            d = np.array([0.0, -0.05, 1.0])
        else:
            target = np.array([rng.uniform(0.2, 0.8) * extent, 0.0, rng.uniform(0.2, 0.8) * extent])
            target[1] = hf.height(target[0], target[2])
            ang = rng.uniform(0, 2 * np.pi)
            off = np.array([np.cos(ang), 0.0, np.sin(ang)]) * rng.uniform(100.0, 300.0)
            eye = target + off + np.array([0.0, rng.uniform(300.0, 700.0), 0.0])
            d = target - eye
        d = d / np.linalg.norm(d)
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.95 else (1.0, 0.0, 0.0)
        # Query boxes: instance bounds nearest the view's line of sight, then
        # props on the terrain around it.
        look = eye + d * 200.0
        dist = np.linalg.norm(centers - look, axis=1)
        near = np.argsort(dist)[: queries_per_view // 2]
        n_props = queries_per_view - near.size
        qx = np.clip(look[0] + rng.uniform(-400, 400, n_props), 0, extent)
        qz = np.clip(look[2] + rng.uniform(-400, 400, n_props), 0, extent)
        qh = np.exp(rng.uniform(math.log(0.25), math.log(3.0), (n_props, 3)))
        qc = np.stack([qx, hf.height(qx, qz) + qh[:, 1], qz], axis=1)
        boxes = np.concatenate([aabb_bounds(centers[near].astype(np.float32),
                                            np.repeat(scale[near, None], 3, axis=1).astype(np.float32)),
                                aabb_bounds(qc.astype(np.float32), qh.astype(np.float32))])
        views.append(dict(size=(w, h), eye=eye.astype(np.float32), dir=d.astype(np.float32), up=up, queries=boxes,
                          mvp=view_projection(eye, d, up, math.pi / 3, w / h)))
    scene.meta.update(views=views, seed=seed, terrain_quads=int(terrain.shape[0]), n_terrain_occluders=n_terrain,
                      instances=n_instances)
    return scene

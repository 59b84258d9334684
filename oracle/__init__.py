"""CPU oracle of the occlusion-culling hot path (ctypes wrapper of oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as the thing measured or shipped.  The product (rasterizer_amd) never
imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# RZ_ORACLE_LIB points the checker at another build of oracle.c (used to
# confirm that the pinning tests fail on a deliberately altered oracle).
LIB_PATH = Path(os.environ["RZ_ORACLE_LIB"]) if os.environ.get("RZ_ORACLE_LIB") else HERE / "_build" / "liboracle.so"

_lib = None


class _Occ(ctypes.Structure):
    _fields_ = [
        ("packets", ctypes.c_void_p),
        ("packetCount", ctypes.c_uint32),
        ("refMin", ctypes.c_float * 4),
        ("refMax", ctypes.c_float * 4),
        ("boundsMin", ctypes.c_float * 4),
        ("boundsMax", ctypes.c_float * 4),
    ]


REF_DIR = Path("/root/reference")
SAH_REF_PATH = HERE / "_ref" / "libsah_ref.so"
QUAD_REF_PATH = HERE / "_ref" / "libquad_ref.so"


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off) and,
    where the reference's sources are present, the reference's own SAH
    batching into oracle/_ref/."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if (REF_DIR / "SoftwareRasterizer" / "SurfaceAreaHeuristic.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REF={REF_DIR}"], check=True)
    return HERE / "_build" / "liboracle.so"


def quad_reference():
    """The reference's QuadDecomposition::decompose as built into oracle/_ref/
    (None if that build is absent): a function (indices, vertices n x 4) ->
    quad indices (4 per quad)."""
    if not QUAD_REF_PATH.exists():
        return None
    lib = ctypes.CDLL(str(QUAD_REF_PATH))
    fn = lib.quad_ref_decompose
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                   ctypes.POINTER(ctypes.c_size_t)]

    def run(indices, vertices):
        idx = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1))
        v = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 4))
        out = np.zeros(max(4 * (idx.size // 3), 1), np.uint32)
        n = ctypes.c_size_t(0)
        if fn(idx.ctypes.data, idx.size, v.ctypes.data, v.shape[0], out.ctypes.data, ctypes.byref(n)) != 0:
            raise ValueError("quad_ref_decompose: index outside the vertex array")
        return out[:n.value].copy()

    return run


def sah_reference():
    """The reference's SurfaceAreaHeuristic::generateBatches as built into
    oracle/_ref/ (None if that build is absent): a function
    (quad_min, quad_max, target, gran) -> list of index arrays, or None where
    the reference's behaviour is undefined."""
    if not SAH_REF_PATH.exists():
        return None
    lib = ctypes.CDLL(str(SAH_REF_PATH))
    fn = lib.sah_ref_batches
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]

    def run(quad_min, quad_max, target, gran):
        mn = np.atleast_2d(np.asarray(quad_min, np.float32))
        mx = np.atleast_2d(np.asarray(quad_max, np.float32))
        n = mn.shape[0]
        boxes = np.zeros((max(n, 1), 8), np.float32)
        boxes[:n, 0:3] = mn[:, :3]
        boxes[:n, 4:7] = mx[:, :3]
        idx = np.zeros(max(n, 1), np.uint32)
        off = np.zeros(n + 2, np.uint32)
        cnt = ctypes.c_uint32(0)
        if fn(n, boxes.ctypes.data, target, gran, idx.ctypes.data, off.ctypes.data, ctypes.byref(cnt)) != 0:
            return None
        return [idx[off[b]:off[b + 1]].astype(np.int64) for b in range(cnt.value)]

    return run


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P, U = ctypes.c_void_p, ctypes.c_uint32
        sig = {
            "or_create": ([U, U], P),
            "or_destroy": ([P], None),
            "or_set_mvp": ([P, P], None),
            "or_clear": ([P], None),
            "or_rasterize": ([P, P, U, P, P, ctypes.c_int], None),
            "or_query": ([P, P, P, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
            "or_query2d": ([P, U, U, U, U, U], ctypes.c_int),
            "or_query_batch": ([P, U, P, P, P], None),
            "or_readback": ([P, P], None),
            "or_get_raw": ([P, P, P], None),
            "or_set_raw": ([P, P, P], None),
            "or_get_table": ([P, P], None),
            "or_get_mvp": ([P, P, P], None),
            "or_get_stats": ([P, P], None),
            "or_reset_stats": ([P], None),
            "or_block_stats": ([P, P], None),
            "or_frame": ([P, U, P, P], None),
            "or_frame_classified": ([P, U, P, P], None),
            "or_frame_labels": ([P, U, P, P, P, P], None),
            "or_setup_dump": ([P, P, U, P, P, ctypes.c_int, P], ctypes.c_int),
            "or_prologue": ([P, P, P, P], None),
            "or_bake": ([P, U, P, P, P, P, P, P], ctypes.c_int),
            "or_bake_order": ([P, U, P], ctypes.c_int),
            "or_x86_rcp": ([ctypes.c_float], ctypes.c_float),
            "or_x86_cvtt": ([ctypes.c_float], ctypes.c_int32),
            "or_bench_frames": ([U, U, U, P, U, P, P, U], ctypes.c_double),
            "or_bench_queries": ([P, U, P, U, U], ctypes.c_double),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _f4(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(4))


def occluder_structs(occluders) -> tuple:
    """ctypes array of OrOccluder for objects with packets / m_refMin / ...
    (keeps the packet arrays alive through the returned tuple)."""
    keep = []
    arr = (_Occ * max(len(occluders), 1))()
    for i, o in enumerate(occluders):
        pk = np.ascontiguousarray(o.packets, dtype=np.uint32)
        keep.append(pk)
        arr[i].packets = pk.ctypes.data
        arr[i].packetCount = pk.size // 8
        for name, src in (("refMin", o.m_refMin), ("refMax", o.m_refMax), ("boundsMin", o.m_boundsMin),
                          ("boundsMax", o.m_boundsMax)):
            getattr(arr[i], name)[:] = [float(x) for x in np.asarray(src, np.float32)]
    return arr, keep


class OracleRasterizer:
    """The reference Rasterizer restated on the CPU (see oracle.h)."""

    def __init__(self, width: int, height: int):
        self._r = lib().or_create(width, height)
        if not self._r:
            raise ValueError("oracle: invalid size or table invariant failure")
        self.width, self.height = width, height
        self.blocks_x, self.blocks_y = width // 8, height // 8

    def __del__(self):
        if getattr(self, "_r", None) and _lib is not None:
            _lib.or_destroy(self._r)
            self._r = None

    def setModelViewProjection(self, m) -> None:
        a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
        lib().or_set_mvp(self._r, _ptr(a))

    def clear(self) -> None:
        lib().or_clear(self._r)

    def rasterize(self, occ, possiblyNearClipped: bool) -> None:
        pk = np.ascontiguousarray(occ.packets, dtype=np.uint32)
        lib().or_rasterize(self._r, _ptr(pk), pk.size // 8, _ptr(_f4(occ.m_refMin)), _ptr(_f4(occ.m_refMax)),
                           1 if possiblyNearClipped else 0)

    def queryVisibility(self, bmin, bmax, needs_clipping: bool = False):
        clip = ctypes.c_int(1 if needs_clipping else 0)
        v = lib().or_query(self._r, _ptr(_f4(bmin)), _ptr(_f4(bmax)), ctypes.byref(clip))
        return bool(v), bool(clip.value)

    def query2D(self, min_x, max_x, min_y, max_y, max_z) -> bool:
        return bool(lib().or_query2d(self._r, min_x, max_x, min_y, max_y, max_z))

    def query_batch(self, bounds: np.ndarray):
        b = np.ascontiguousarray(np.asarray(bounds, np.float32).reshape(-1, 8))
        vis = np.zeros(max(b.shape[0], 1), np.uint8)
        clip = np.zeros(max(b.shape[0], 1), np.uint8)
        lib().or_query_batch(self._r, b.shape[0], _ptr(b), _ptr(vis), _ptr(clip))
        return vis[: b.shape[0]].astype(bool), clip[: b.shape[0]].astype(bool)

    def readBackDepth(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        lib().or_readback(self._r, _ptr(out))
        return out

    def raw(self):
        depth = np.zeros(self.blocks_x * self.blocks_y * 64, np.uint16)
        hiz = np.zeros(self.blocks_x * self.blocks_y + 8, np.uint16)
        lib().or_get_raw(self._r, _ptr(depth), _ptr(hiz))
        return depth, hiz

    def set_raw(self, depth=None, hiz=None) -> None:
        d = None if depth is None else np.ascontiguousarray(depth, np.uint16)
        h = None if hiz is None else np.ascontiguousarray(hiz, np.uint16)
        lib().or_set_raw(self._r, _ptr(d) if d is not None else None, _ptr(h) if h is not None else None)

    def table(self) -> np.ndarray:
        t = np.zeros(4096, np.uint64)
        lib().or_get_table(self._r, _ptr(t))
        return t

    def mvp_state(self):
        pre = np.zeros(16, np.float32)
        raw = np.zeros(16, np.float32)
        lib().or_get_mvp(self._r, _ptr(pre), _ptr(raw))
        return pre, raw

    def stats(self) -> dict:
        s = np.zeros(7, np.uint64)
        lib().or_get_stats(self._r, _ptr(s))
        keys = ["visited", "hiz_rejected", "offset_rejected", "empty_mask", "written", "quads_setup", "quads_valid"]
        return dict(zip(keys, (int(x) for x in s)))

    def reset_stats(self) -> None:
        lib().or_reset_stats(self._r)

    def block_stats(self, enable_only: bool = False):
        """Per-block (visited, written) counts since the last reset, shape
        (blocksY, blocksX, 2); the first call switches counting on."""
        if enable_only:
            lib().or_block_stats(self._r, None)
            return None
        out = np.zeros((self.blocks_y, self.blocks_x, 2), np.uint32)
        lib().or_block_stats(self._r, _ptr(out))
        return out

    def setup_dump(self, occ, clip: bool) -> np.ndarray:
        """Per-quad setup of rasterize<clip> for one occluder (24 floats per
        quad: valid, mode, minX, minY, rangeX, rangeY, maxZ, ...)."""
        pk = np.ascontiguousarray(occ.packets, dtype=np.uint32)
        out = np.zeros((pk.size // 4, 24), np.float32)
        n = lib().or_setup_dump(self._r, _ptr(pk), pk.size // 8, _ptr(_f4(occ.m_refMin)), _ptr(_f4(occ.m_refMax)),
                                1 if clip else 0, _ptr(out))
        return out[:n]

    def frame_labels(self, occluders, order=None):
        """Labelled classify-then-draw (oracle.h or_frame_labels): returns
        (per-occluder status, per-batch label counts)."""
        arr, keep = occluder_structs(occluders)
        nb = sum(int(np.asarray(o.packets).size) // 32 for o in occluders)
        status = np.zeros(max(len(occluders), 1), np.uint8)
        out = np.zeros(max(nb, 1), np.uint32)
        ordp = None
        if order is not None:
            order = np.ascontiguousarray(order, np.uint32)
            if sorted(order.tolist()) != list(range(nb)):
                raise ValueError("order must be a permutation of the batch numbers")
            ordp = _ptr(order)
        lib().or_frame_labels(self._r, len(occluders), ctypes.cast(arr, ctypes.c_void_p), ordp, _ptr(status),
                              _ptr(out))
        del keep
        return status[: len(occluders)], out[:nb]

    def prologue(self, occ) -> np.ndarray:
        """or_prologue: T (16, row-major) then c0, c1 for an occluder."""
        out = np.zeros(18, np.float32)
        lib().or_prologue(self._r, _ptr(_f4(occ.m_refMin)), _ptr(_f4(occ.m_refMax)), _ptr(out))
        return out

    def draw_frame(self, occluders, gated: bool = True) -> np.ndarray:
        """Main.cpp:192-206 (gated) or the batched path's classify-then-draw
        sequence (gated=False); returns per-occluder status 0/1/2."""
        arr, keep = occluder_structs(occluders)
        status = np.zeros(max(len(occluders), 1), np.uint8)
        fn = lib().or_frame if gated else lib().or_frame_classified
        fn(self._r, len(occluders), ctypes.cast(arr, ctypes.c_void_p), _ptr(status))
        del keep
        return status[: len(occluders)]


def bake(vertices, ref_min, ref_max):
    """Occluder::bake restated (oracle.c); returns packets, boundsMin, boundsMax, center."""
    v = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 4))
    n = v.shape[0]
    packets = np.zeros(n, np.uint32)
    bmin, bmax, cen = (np.zeros(4, np.float32) for _ in range(3))
    rc = lib().or_bake(_ptr(v), n, _ptr(_f4(ref_min)), _ptr(_f4(ref_max)), _ptr(packets), _ptr(bmin), _ptr(bmax),
                       _ptr(cen))
    if rc != 0:
        raise ValueError("oracle bake: vertex count must be a positive multiple of 32")
    return packets, bmin, bmax, cen


def bake_order(vertices) -> np.ndarray:
    """The quad order of Occluder::bake's k-means regroup (Occluder.cpp:11-93), restated (oracle.c)."""
    v = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 4))
    nq = v.shape[0] // 4
    out = np.zeros(nq, np.uint32)
    if nq == 0 or lib().or_bake_order(_ptr(v), nq, _ptr(out)) != 0:
        raise ValueError("oracle bake_order: needs at least one quad")
    return out


def x86_rcp(x: float) -> float:
    return float(lib().or_x86_rcp(ctypes.c_float(x)))


def x86_cvtt(x: float) -> int:
    return int(lib().or_x86_cvtt(ctypes.c_float(x)))


def bench_frames(width, height, occluders, mvps, orders, threads: int) -> float:
    """Wall seconds for len(mvps) frames on `threads` threads (one oracle per thread)."""
    arr, keep = occluder_structs(occluders)
    m = np.ascontiguousarray(np.asarray(mvps, np.float32).reshape(-1, 16))
    o = np.ascontiguousarray(np.asarray(orders, np.uint32).reshape(m.shape[0], len(occluders)))
    t = lib().or_bench_frames(width, height, len(occluders), ctypes.cast(arr, ctypes.c_void_p), m.shape[0], _ptr(m),
                              _ptr(o), threads)
    del keep
    return float(t)


def bench_queries(rast: OracleRasterizer, bounds: np.ndarray, reps: int, threads: int) -> float:
    b = np.ascontiguousarray(np.asarray(bounds, np.float32).reshape(-1, 8))
    return float(lib().or_bench_queries(rast._r, b.shape[0], _ptr(b), reps, threads))

"""MI355X occlusion-culling rasterizer: Python binding of the C ABI.

The classes mirror the reference's C++ interface (rawrunprotected/rasterizer,
SoftwareRasterizer/Rasterizer.h:10-61 and Occluder.h:7-21) so that tests and
benchmarks read like calls into the reference.  All work runs in
``_lib/librasterizer_amd.so`` (hand-written gfx950 kernels); there is no CPU
fallback: if the library or a GPU is missing, construction raises.
"""

from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "_lib" / "librasterizer_amd.so"

RZ_OK = 0
_STATUS = {
    -1: "invalid argument",
    -2: "out of memory",
    -3: "HIP runtime or device error",
    -4: "device-side wait timed out",
    -5: "host CPU RCPPS has an unsupported structure",
    -6: "buffer too large for one co-resident frame launch",
    -7: "launch configurations disagree (autotuner check)",
    -8: "file cannot be written or read, or is not this format, version or checksum",
}


class RasterizerError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {_STATUS.get(status, 'unknown status')} ({status})")
        self.status = status


_lib = None

# Objects holding device memory, released in order at interpreter exit (before
# C-level exit handlers run): buffers and occluder copies, then batches and
# models, then contexts.
_live = [weakref.WeakSet() for _ in range(3)]


def _release_all() -> None:
    for tier in _live:
        for obj in list(tier):
            try:
                (getattr(obj, "close", None) or getattr(obj, "release", None) or getattr(obj, "free"))()
            except Exception:
                pass


atexit.register(_release_all)

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I32P = ctypes.POINTER(ctypes.c_int)

_SIGNATURES = {
    "rz_status_string": ([ctypes.c_int], ctypes.c_char_p),
    "rz_create": ([_U32, _U32, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_destroy": ([_P], ctypes.c_int),
    "rz_set_mvp": ([_P, _P], ctypes.c_int),
    "rz_clear": ([_P], ctypes.c_int),
    "rz_occluder_create": ([_P, _P, _U32, _P, _P, _P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_occluder_create_device": ([_P, _P, _U32, _P, _P, _P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_occluder_destroy": ([_P], ctypes.c_int),
    "rz_rasterize": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "rz_draw_frame": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_draw_frame_noalloc": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_set_wait_timeout": ([_P, ctypes.c_uint64], ctypes.c_int),
    "rz_query": ([_P, _P, _P, _I32P, _I32P], ctypes.c_int),
    "rz_query2d": ([_P, _U32, _U32, _U32, _U32, _U32, _I32P], ctypes.c_int),
    "rz_query_batch": ([_P, _U32, _P, _P, _P], ctypes.c_int),
    "rz_query_batch_device": ([_P, _U32, _P, _P, _P], ctypes.c_int),
    "rz_readback_depth": ([_P, _P], ctypes.c_int),
    "rz_readback_raw": ([_P, _P, _P], ctypes.c_int),
    "rz_upload_raw": ([_P, _P, _P], ctypes.c_int),
    "rz_sync": ([_P], ctypes.c_int),
    "rz_get_stream": ([_P, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_get_info": ([_P, _P], ctypes.c_int),
    "rz_get_tables": ([_P, _P, _P], ctypes.c_int),
    "rz_get_mvp": ([_P, _P, _P], ctypes.c_int),
    "rz_bake": ([_P, _U32, _P, _P, _P, _P, _P, _P], ctypes.c_int),
    "rz_bake_device": ([_P, _U32, _P, _P, _P, _P, _P], ctypes.c_int),
    "rz_sah_batches": ([_P, _U32, _P, _U32, _U32, _P, _P, ctypes.POINTER(_U32)], ctypes.c_int),
    "rz_prepare_scene_device": ([_P, _U32, _P, _U32, _U32, _P, _P, _P, ctypes.POINTER(_U32), _P], ctypes.c_int),
    "rz_timer_start": ([_P], ctypes.c_int),
    "rz_timer_stop": ([_P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "rz_profile_enable": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_profile_read": ([_P, _P, _P], ctypes.c_int),
    "rz_frame_stats_enable": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_frame_stats_read": ([_P, _P, ctypes.POINTER(_U32)], ctypes.c_int),
    "rz_device_alloc": ([_P, ctypes.c_size_t, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_device_read": ([_P, _P, ctypes.c_size_t, _P], ctypes.c_int),
    "rz_device_free": ([_P], ctypes.c_int),
    "rz_views_create": ([_U32, _P, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_views_destroy": ([_P], ctypes.c_int),
    "rz_views_set_mvp": ([_P, _U32, _P], ctypes.c_int),
    "rz_views_set_occluders": ([_P, _U32, _U32, _P], ctypes.c_int),
    "rz_views_set_queries": ([_P, _U32, _U32, _P], ctypes.c_int),
    "rz_views_run": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_views_sync": ([_P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "rz_views_read_raw": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_views_read_queries": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_views_read_status": ([_P, _U32, _P], ctypes.c_int),
    "rz_views_get_info": ([_P, _P], ctypes.c_int),
    "rz_views_enable_labels": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_views_set_batch_order": ([_P, _U32, _U32, _P], ctypes.c_int),
    "rz_views_read_labels": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_mlp_create": ([_U32, ctypes.POINTER(_P)], ctypes.c_int),
    "rz_mlp_destroy": ([_P], ctypes.c_int),
    "rz_mlp_set_params": ([_P, _P, _P], ctypes.c_int),
    "rz_mlp_set_uses_labels": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_mlp_get_params": ([_P, _P, _P], ctypes.c_int),
    "rz_mlp_forward": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_mlp_gradients": ([_P, _U32, _P, _U32, _P, _P, _P], ctypes.c_int),
    "rz_mlp_train_step": ([_P, _U32, _P, _U32, _P, ctypes.c_float, _P], ctypes.c_int),
    "rz_views_set_model": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "rz_views_read_order": ([_P, _U32, _P], ctypes.c_int),
    "rz_views_enable_stats": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_views_read_stats": ([_P, _P], ctypes.c_int),
    "rz_views_read_tile_stats": ([_P, _P], ctypes.c_int),
    "rz_views_set_eye": ([_P, ctypes.c_uint32, _P], ctypes.c_int),
    "rz_get_config": ([_P, _P], ctypes.c_int),
    "rz_set_config": ([_P, _U32, _U32], ctypes.c_int),
    "rz_tune_frame": ([_P, _U32, _P, _U32, ctypes.c_char_p, _P, _P], ctypes.c_int),
    "rz_load_config": ([_P, ctypes.c_char_p, _P], ctypes.c_int),
    "rz_views_get_config": ([_P, _P], ctypes.c_int),
    "rz_views_set_config": ([_P, _U32, _U32, _U32], ctypes.c_int),
    "rz_views_tune": ([_P, _U32, ctypes.c_char_p, _P, _P], ctypes.c_int),
    "rz_views_load_config": ([_P, ctypes.c_char_p, _P], ctypes.c_int),
    "rz_views_set_query_gated": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_views_set_wait_timeout": ([_P, ctypes.c_uint64], ctypes.c_int),
    "rz_views_profile_enable": ([_P, ctypes.c_int], ctypes.c_int),
    "rz_views_profile_read": ([_P, _P], ctypes.c_int),
    "rz_views_read_items": ([_P, _U32, _P, _P], ctypes.c_int),
    "rz_quad_decompose": ([_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.POINTER(ctypes.c_size_t)],
                          ctypes.c_int),
    "rz_quad_decompose_device": ([_P, _U32, _P, _P, _U32, _P, ctypes.POINTER(_U32)], ctypes.c_int),
    "rz_prepare_mesh_device": ([_P, _U32, _P, _P, _U32, _U32, _U32, _P, _P, _P, ctypes.POINTER(_U32), _P],
                               ctypes.c_int),
    "rz_write_occluders": ([ctypes.c_char_p, _U32, _P], ctypes.c_int),
    "rz_read_occluders": ([ctypes.c_char_p, _P, _P, _P, _P], ctypes.c_int),
    "rz_write_frame": ([ctypes.c_char_p, _U32, _U32, _P, _P, _U32, _P, _P], ctypes.c_int),
    "rz_read_frame": ([ctypes.c_char_p, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
}


class _OccluderData(ctypes.Structure):
    """rz_occluder_data: a baked occluder's packets, reference box and bounds."""
    _fields_ = [("packets", ctypes.POINTER(ctypes.c_uint32)), ("packetCount", ctypes.c_uint32),
                ("refMin", ctypes.c_float * 4), ("refMax", ctypes.c_float * 4),
                ("boundsMin", ctypes.c_float * 4), ("boundsMax", ctypes.c_float * 4)]


class FrameConfig(ctypes.Structure):
    """rz_frame_config: tile edge in 8x8 blocks, wave64s per tile, tuned ms."""
    _fields_ = [("tileW", ctypes.c_uint32), ("waves", ctypes.c_uint32), ("ms", ctypes.c_float)]


class ViewsConfig(ctypes.Structure):
    """rz_views_config: tile edge in 8x8 blocks, wave64s per work-group, work-groups per CU (0: all that
    fit), tuned ms."""
    _fields_ = [("tileW", ctypes.c_uint32), ("waves", ctypes.c_uint32), ("groupsPerCU", ctypes.c_uint32),
                ("ms", ctypes.c_float)]


def _cfg_dict(c) -> dict:
    return {name: (float(getattr(c, name)) if name == "ms" else int(getattr(c, name))) for name, _ in c._fields_}


def lib() -> ctypes.CDLL:
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m rasterizer_amd.build` "
                "(the rasterizer has no CPU fallback)")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def _check(status: int, what: str) -> None:
    if status != RZ_OK:
        raise RasterizerError(status, what)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _vec4(v) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(4))
    return a


class Occluder:
    """Baked occluder (Occluder.h:7-21): packets plus its reference box and bounds."""

    def __init__(self, packets: np.ndarray, ref_min, ref_max, bounds_min, bounds_max, center=None):
        self.packets = np.ascontiguousarray(packets, dtype=np.uint32)
        if self.packets.size % 32:
            raise ValueError("packets must hold whole 8-quad batches (32 words each)")
        self.m_refMin = _vec4(ref_min)
        self.m_refMax = _vec4(ref_max)
        self.m_boundsMin = _vec4(bounds_min)
        self.m_boundsMax = _vec4(bounds_max)
        self.m_center = _vec4(center if center is not None else (self.m_boundsMin + self.m_boundsMax) * 0.5)
        self._device = {}
        _live[0].add(self)

    @property
    def m_packetCount(self) -> int:
        return self.packets.size // 8

    @property
    def quad_count(self) -> int:
        return self.packets.size // 4

    @staticmethod
    def bake(vertices, ref_min, ref_max) -> "Occluder":
        """Occluder::bake (Occluder.cpp:7-181), run by the library's host code."""
        v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 4))
        n = v.shape[0]
        if n == 0 or n % 32:
            raise ValueError("vertex count must be a positive multiple of 32 (whole 8-quad batches)")
        packets = np.zeros(n, dtype=np.uint32)
        lo, hi = _vec4(ref_min), _vec4(ref_max)
        bmin, bmax, cen = (np.zeros(4, np.float32) for _ in range(3))
        _check(lib().rz_bake(_ptr(v), n, _ptr(lo), _ptr(hi), _ptr(packets), _ptr(bmin), _ptr(bmax), _ptr(cen)),
               "Occluder.bake")
        return Occluder(packets, lo, hi, bmin, bmax, cen)

    def device(self, ctx) -> ctypes.c_void_p:
        key = ctx.value
        h = self._device.get(key)
        if h is None:
            h = ctypes.c_void_p()
            _check(lib().rz_occluder_create(ctx, _ptr(self.packets), self.m_packetCount, _ptr(self.m_refMin),
                                            _ptr(self.m_refMax), _ptr(self.m_boundsMin), _ptr(self.m_boundsMax),
                                            ctypes.byref(h)), "rz_occluder_create")
            self._device[key] = h
        return h

    def release(self) -> None:
        if _lib is not None:
            for h in self._device.values():
                _lib.rz_occluder_destroy(h)
        self._device = {}

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def save_occluders(path, occluders) -> None:
    """Writes baked occluders (packets, reference boxes, bounds) to a
    versioned, checksummed file (rz_write_occluders)."""
    occluders = list(occluders)
    arr = (_OccluderData * max(len(occluders), 1))()
    for d, o in zip(arr, occluders):
        d.packets = o.packets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        d.packetCount = o.packets.size // 8
        for name, v in (("refMin", o.m_refMin), ("refMax", o.m_refMax), ("boundsMin", o.m_boundsMin),
                        ("boundsMax", o.m_boundsMax)):
            getattr(d, name)[:] = [float(x) for x in v]
    _check(lib().rz_write_occluders(str(path).encode(), len(occluders), ctypes.cast(arr, _P)), "rz_write_occluders")


def load_occluders(path) -> list:
    """Reads a file written by save_occluders into Occluder objects."""
    n, total = ctypes.c_uint32(0), ctypes.c_uint32(0)
    p = str(path).encode()
    _check(lib().rz_read_occluders(p, ctypes.byref(n), ctypes.byref(total), None, None), "rz_read_occluders")
    arr = (_OccluderData * max(n.value, 1))()
    words = np.zeros(max(total.value * 8, 1), np.uint32)
    _check(lib().rz_read_occluders(p, ctypes.byref(n), ctypes.byref(total), ctypes.cast(arr, _P), _ptr(words)),
           "rz_read_occluders")
    out, at = [], 0
    for d in arr[: n.value]:
        k = d.packetCount * 8
        out.append(Occluder(words[at:at + k].copy(), list(d.refMin), list(d.refMax), list(d.boundsMin),
                            list(d.boundsMax)))
        at += k
    return out


def save_frame(path, width: int, height: int, depth: np.ndarray, hiz: np.ndarray, visible=None, clip=None) -> None:
    """Writes a frame dump: raw depth and hiZ (as Rasterizer.raw returns
    them) and optionally the query bits (rz_write_frame)."""
    depth = np.ascontiguousarray(depth, np.uint16)
    hiz = np.ascontiguousarray(hiz, np.uint16)
    nb = (width // 8) * (height // 8)
    if depth.size != nb * 64 or hiz.size != nb + 8:
        raise ValueError("depth must hold (w/8)*(h/8)*64 and hiz (w/8)*(h/8) + 8 values")
    nq = 0 if visible is None else len(visible)
    vis = np.ascontiguousarray(visible if visible is not None else np.zeros(1), np.uint8)
    clp = np.ascontiguousarray(clip if clip is not None else np.zeros(max(nq, 1)), np.uint8)
    if clp.size < nq:
        raise ValueError("clip must have one byte per query")
    _check(lib().rz_write_frame(str(path).encode(), width, height, _ptr(depth), _ptr(hiz), nq, _ptr(vis), _ptr(clp)),
           "rz_write_frame")


def load_frame(path) -> dict:
    """Reads a frame dump: {width, height, depth, hiz, visible, clip}."""
    w, h, nq = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    p = str(path).encode()
    _check(lib().rz_read_frame(p, ctypes.byref(w), ctypes.byref(h), ctypes.byref(nq), None, None, None, None),
           "rz_read_frame")
    nb = (w.value // 8) * (h.value // 8)
    depth = np.zeros(nb * 64, np.uint16)
    hiz = np.zeros(nb + 8, np.uint16)
    vis = np.zeros(max(nq.value, 1), np.uint8)
    clp = np.zeros(max(nq.value, 1), np.uint8)
    _check(lib().rz_read_frame(p, ctypes.byref(w), ctypes.byref(h), ctypes.byref(nq), _ptr(depth), _ptr(hiz),
                               _ptr(vis), _ptr(clp)), "rz_read_frame")
    return {"width": w.value, "height": h.value, "depth": depth, "hiz": hiz, "visible": vis[: nq.value],
            "clip": clp[: nq.value]}


class DeviceBuffer:
    """Device memory owned through the library (for device-pointer entry points)."""

    def __init__(self, rast: "Rasterizer", nbytes: int, init: np.ndarray | None = None):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        src = None
        if init is not None:
            self._init = np.ascontiguousarray(init)
            if self._init.nbytes != self.nbytes:
                raise ValueError("init size does not match nbytes")
            src = _ptr(self._init)
        _check(lib().rz_device_alloc(rast.ctx, self.nbytes, src, ctypes.byref(self.ptr)), "rz_device_alloc")
        self._rast = rast
        _live[0].add(self)

    def read(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(count, dtype)
        _check(lib().rz_device_read(self._rast.ctx, self.ptr, out.nbytes, _ptr(out)), "rz_device_read")
        return out

    def free(self) -> None:
        if self.ptr and _lib is not None:
            _lib.rz_device_free(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Rasterizer:
    """One depth buffer + hiZ on the GPU (Rasterizer.h:10-61)."""

    def __init__(self, width: int, height: int):
        self._ctx = ctypes.c_void_p()
        _check(lib().rz_create(width, height, ctypes.byref(self._ctx)), "Rasterizer")
        self.width, self.height = width, height
        self.blocks_x, self.blocks_y = width // 8, height // 8
        info = np.zeros(6, np.uint32)
        _check(lib().rz_get_info(self._ctx, _ptr(info)), "rz_get_info")
        self.tile = (int(info[4]), int(info[5]))
        _live[2].add(self)

    @property
    def ctx(self) -> ctypes.c_void_p:
        return self._ctx

    def close(self) -> None:
        if self._ctx and _lib is not None:
            _lib.rz_destroy(self._ctx)
        self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setModelViewProjection(self, matrix) -> None:
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.float32).reshape(16))
        _check(lib().rz_set_mvp(self._ctx, _ptr(m)), "setModelViewProjection")

    def clear(self) -> None:
        _check(lib().rz_clear(self._ctx), "clear")

    def rasterize(self, occluder: Occluder, possiblyNearClipped: bool) -> None:
        _check(lib().rz_rasterize(self._ctx, occluder.device(self._ctx), 1 if possiblyNearClipped else 0),
               "rasterize")

    def queryVisibility(self, bounds_min, bounds_max, needs_clipping: bool = False):
        """Returns (visible, needsClipping); needsClipping keeps the passed value
        when the box is frustum-culled, as the reference leaves it unwritten."""
        lo, hi = _vec4(bounds_min), _vec4(bounds_max)
        vis = ctypes.c_int(0)
        clip = ctypes.c_int(1 if needs_clipping else 0)
        _check(lib().rz_query(self._ctx, _ptr(lo), _ptr(hi), ctypes.byref(vis), ctypes.byref(clip)),
               "queryVisibility")
        return bool(vis.value), bool(clip.value)

    def query2D(self, min_x, max_x, min_y, max_y, max_z) -> bool:
        vis = ctypes.c_int(0)
        _check(lib().rz_query2d(self._ctx, min_x, max_x, min_y, max_y, max_z, ctypes.byref(vis)), "query2D")
        return bool(vis.value)

    def readBackDepth(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        _check(lib().rz_readback_depth(self._ctx, _ptr(out)), "readBackDepth")
        return out

    # Extensions beyond the reference class.
    def draw_frame(self, occluders, want_status: bool = True):
        """Main.cpp:192-206's query-gated loop on the device; returns per-occluder
        status (0 not drawn, 1 rasterize<false>, 2 rasterize<true>)."""
        n = len(occluders)
        handles = (ctypes.c_void_p * max(n, 1))(*[o.device(self._ctx).value for o in occluders])
        status = np.zeros(max(n, 1), np.uint8)
        _check(lib().rz_draw_frame(self._ctx, n, ctypes.cast(handles, ctypes.c_void_p),
                                   _ptr(status) if want_status else None), "draw_frame")
        if want_status:
            self.sync()
        return status[:n]

    def draw_frame_handles(self, handles, n: int) -> None:
        """rz_draw_frame on a prepared ctypes array of device handles (no status)."""
        _check(lib().rz_draw_frame(self._ctx, n, ctypes.cast(handles, ctypes.c_void_p), None), "draw_frame")

    def set_wait_timeout(self, ticks: int) -> None:
        """Bound of each cross-tile wait of draw_frame in 100 MHz ticks (0: the default 2 s)."""
        _check(lib().rz_set_wait_timeout(self._ctx, int(ticks)), "rz_set_wait_timeout")

    def handle_array(self, occluders):
        return (ctypes.c_void_p * max(len(occluders), 1))(*[o.device(self._ctx).value for o in occluders])

    def config(self) -> dict:
        c = FrameConfig()
        _check(lib().rz_get_config(self._ctx, ctypes.byref(c)), "rz_get_config")
        return {"tileW": c.tileW, "waves": c.waves}

    def set_config(self, tileW: int, waves: int) -> None:
        _check(lib().rz_set_config(self._ctx, tileW, waves), "rz_set_config")

    def autotune(self, occluders, reps: int = 5, cache: "str | os.PathLike | None" = None) -> dict:
        """rz_tune_frame: times every frame-kernel configuration on this frame
        (with identical bits required) and keeps the fastest."""
        n = len(occluders)
        handles = self.handle_array(occluders)
        cands = (FrameConfig * 6)()
        k = ctypes.c_uint32(0)
        _check(lib().rz_tune_frame(self._ctx, n, ctypes.cast(handles, ctypes.c_void_p), reps,
                                   os.fsencode(cache) if cache is not None else None, cands, ctypes.byref(k)),
               "rz_tune_frame")
        return {"chosen": self.config(), "candidates": [_cfg_dict(cands[i]) for i in range(k.value)]}

    def load_config(self, cache) -> bool:
        found = ctypes.c_int(0)
        _check(lib().rz_load_config(self._ctx, os.fsencode(cache), ctypes.byref(found)), "rz_load_config")
        return bool(found.value)

    def bake_device(self, meshes, ref_mins, ref_maxs) -> list:
        """Occluder::bake on the device for several occluders at once
        (rz_bake_device): meshes[i] is an (n_i, 4) vertex array (n_i a
        positive multiple of 32), with its reference box; returns Occluder
        objects, bit-identical to Occluder.bake."""
        meshes = [np.ascontiguousarray(np.asarray(m, np.float32).reshape(-1, 4)) for m in meshes]
        n = len(meshes)
        if n == 0:
            return []
        offsets = np.zeros(n + 1, np.uint32)
        offsets[1:] = np.cumsum([m.shape[0] for m in meshes])
        verts = np.concatenate(meshes)
        ref = np.concatenate([np.concatenate([_vec4(a), _vec4(b)]) for a, b in zip(ref_mins, ref_maxs)]).astype(np.float32)
        d_v = DeviceBuffer(self, verts.nbytes, verts)
        d_p = DeviceBuffer(self, int(offsets[-1]) * 4)
        d_b = DeviceBuffer(self, n * 48)
        try:
            _check(lib().rz_bake_device(self.ctx, n, _ptr(offsets), d_v.ptr, _ptr(ref), d_p.ptr, d_b.ptr),
                   "rz_bake_device")
            packets = d_p.read(np.uint32, int(offsets[-1]))
            bounds = d_b.read(np.float32, n * 12).reshape(n, 12)
        finally:
            for b in (d_v, d_p, d_b):
                b.free()
        return [Occluder(packets[offsets[i]:offsets[i + 1]].copy(), ref[8 * i:8 * i + 4], ref[8 * i + 4:8 * i + 8],
                         bounds[i, 0:4], bounds[i, 4:8], bounds[i, 8:12]) for i in range(n)]

    def sah_batches_device(self, quad_min, quad_max, target: int = 512, gran: int = 8) -> list:
        """SurfaceAreaHeuristic::generateBatches on the device
        (rz_sah_batches): quad_min and quad_max are (n, 3) or (n, 4) box
        corners; returns the batches' quad-index arrays in the reference's
        order, identical to scenes.sah_batches."""
        mn = np.atleast_2d(np.asarray(quad_min, np.float32))
        mx = np.atleast_2d(np.asarray(quad_max, np.float32))
        n = mn.shape[0]
        boxes = np.zeros((n, 8), np.float32)
        boxes[:, 0:3] = mn[:, :3]
        boxes[:, 4:7] = mx[:, :3]
        offsets = np.zeros(n + 1, np.uint32)
        count = _U32(0)
        if n == 0:
            _check(lib().rz_sah_batches(self.ctx, 0, None, target, gran, None, _ptr(offsets), ctypes.byref(count)),
                   "rz_sah_batches")
            return []
        d_b = DeviceBuffer(self, boxes.nbytes, boxes)
        d_i = DeviceBuffer(self, n * 4)
        try:
            _check(lib().rz_sah_batches(self.ctx, n, d_b.ptr, target, gran, d_i.ptr, _ptr(offsets),
                                        ctypes.byref(count)), "rz_sah_batches")
            idx = d_i.read(np.uint32, n).astype(np.int64)
        finally:
            d_b.free()
            d_i.free()
        return [idx[offsets[k]:offsets[k + 1]] for k in range(count.value)]

    def bake_scene_device(self, quads, target: int = 512, gran: int = 8) -> list:
        """Main.cpp:86-128 with both scene-preparation steps on the device:
        pad to whole 8-quad batches, SAH-batch the quads (rz_sah_batches),
        then bake every batch against the scene box (rz_bake_device).
        quads: (n, 4, 4) vertices.  Returns the Occluders, identical to
        scenes.bake_scene(quads).occluders."""
        from . import scenes

        q = scenes.pad_to_batches(np.asarray(quads, np.float32))
        verts = q.reshape(-1, 4)
        ref_min = verts.min(axis=0).astype(np.float32)
        ref_max = verts.max(axis=0).astype(np.float32)
        batches = self.sah_batches_device(q.min(axis=1), q.max(axis=1), target, gran)
        meshes = [q[b].reshape(-1, 4) for b in batches]
        return self.bake_device(meshes, [ref_min] * len(meshes), [ref_max] * len(meshes))

    def prepare_scene_device(self, quads, target: int = 512, gran: int = 8) -> list:
        """rz_prepare_scene_device: Main.cpp:86-128 entirely on the device
        from (n, 4, 4) quad vertices uploaded once (padding, SAH batching,
        scene box, bake).  Returns the Occluders, identical to
        scenes.bake_scene(quads).occluders."""
        q = np.ascontiguousarray(np.asarray(quads, np.float32).reshape(-1, 4, 4))
        n = q.shape[0]
        if n == 0:
            offsets = np.zeros(1, np.uint32)
            ref = np.zeros(8, np.float32)
            count = _U32(0)
            _check(lib().rz_prepare_scene_device(self.ctx, 0, None, target, gran, None, None, _ptr(offsets),
                                                 ctypes.byref(count), _ptr(ref)), "rz_prepare_scene_device")
            return []
        n_pad = (n + 7) // 8 * 8
        max_occ = n_pad // 8
        offsets = np.zeros(max_occ + 1, np.uint32)
        ref = np.zeros(8, np.float32)
        count = _U32(0)
        d_v = DeviceBuffer(self, q.nbytes, q)
        d_p = DeviceBuffer(self, n_pad * 16)
        d_b = DeviceBuffer(self, max_occ * 48)
        try:
            _check(lib().rz_prepare_scene_device(self.ctx, n, d_v.ptr, target, gran, d_p.ptr, d_b.ptr, _ptr(offsets),
                                                 ctypes.byref(count), _ptr(ref)), "rz_prepare_scene_device")
            k = count.value
            packets = d_p.read(np.uint32, n_pad * 4)
            bounds = d_b.read(np.float32, max_occ * 12).reshape(max_occ, 12)
        finally:
            for b in (d_v, d_p, d_b):
                b.free()
        return [Occluder(packets[4 * offsets[i]:4 * offsets[i + 1]].copy(), ref[0:4], ref[4:8], bounds[i, 0:4],
                         bounds[i, 4:8], bounds[i, 8:12]) for i in range(k)]

    def prepare_mesh_device(self, indices, vertices, target: int = 512, gran: int = 8) -> list:
        """rz_prepare_mesh_device: Main.cpp:86-128 on the device from an
        indexed triangle mesh uploaded once (QuadDecomposition, padding, SAH
        batching, the vertex buffer's box, bake).  Returns the Occluders,
        identical to scenes.bake_scene(scenes.decompose_mesh(indices,
        vertices), vertex_buffer=vertices).occluders."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
        n_tri = idx.size // 3
        idx = np.ascontiguousarray(idx[:3 * n_tri])
        v = np.asarray(vertices, dtype=np.float32)
        if v.ndim == 2 and v.shape[1] == 3:
            v = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
        v = np.ascontiguousarray(v.reshape(-1, 4))
        max_occ = (n_tri + 7) // 8
        offsets = np.zeros(max_occ + 1, np.uint32)
        ref = np.zeros(8, np.float32)
        count = _U32(0)
        if n_tri == 0:
            _check(lib().rz_prepare_mesh_device(self.ctx, 0, None, None, v.shape[0], target, gran, None, None,
                                                _ptr(offsets), ctypes.byref(count), _ptr(ref)),
                   "rz_prepare_mesh_device")
            return []
        d_i = DeviceBuffer(self, idx.nbytes, idx)
        d_v = DeviceBuffer(self, max(v.nbytes, 16), v if v.nbytes else None)
        d_p = DeviceBuffer(self, max_occ * 8 * 16)
        d_b = DeviceBuffer(self, max_occ * 48)
        try:
            _check(lib().rz_prepare_mesh_device(self.ctx, n_tri, d_i.ptr, d_v.ptr, v.shape[0], target, gran, d_p.ptr,
                                                d_b.ptr, _ptr(offsets), ctypes.byref(count), _ptr(ref)),
                   "rz_prepare_mesh_device")
            k = count.value
            packets = d_p.read(np.uint32, max_occ * 32)
            bounds = d_b.read(np.float32, max_occ * 12).reshape(max_occ, 12)
        finally:
            for b in (d_i, d_v, d_p, d_b):
                b.free()
        return [Occluder(packets[4 * offsets[i]:4 * offsets[i + 1]].copy(), ref[0:4], ref[4:8], bounds[i, 0:4],
                         bounds[i, 4:8], bounds[i, 8:12]) for i in range(k)]

    def quad_decompose_device(self, indices, vertices) -> np.ndarray:
        """rz_quad_decompose_device: QuadDecomposition::decompose with the
        candidate graph and the quads built on the GPU (the matching on the
        host) from a mesh uploaded once; the same quads as
        QuadDecomposition.decompose."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
        n_tri = idx.size // 3
        idx = idx[:3 * n_tri]
        v = np.asarray(vertices, dtype=np.float32)
        if v.ndim == 2 and v.shape[1] == 3:
            v = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
        v = np.ascontiguousarray(v.reshape(-1, 4))
        n = _U32(0)
        if n_tri == 0:
            _check(lib().rz_quad_decompose_device(self.ctx, 0, None, None, v.shape[0], None, ctypes.byref(n)),
                   "rz_quad_decompose_device")
            return np.zeros(0, np.uint32)
        d_i = DeviceBuffer(self, idx.nbytes, idx)
        d_v = DeviceBuffer(self, max(v.nbytes, 16), v if v.nbytes else None)
        d_o = DeviceBuffer(self, 16 * n_tri)
        try:
            _check(lib().rz_quad_decompose_device(self.ctx, n_tri, d_i.ptr, d_v.ptr, v.shape[0], d_o.ptr,
                                                  ctypes.byref(n)), "rz_quad_decompose_device")
            out = d_o.read(np.uint32, n.value) if n.value else np.zeros(0, np.uint32)
        finally:
            for b in (d_i, d_v, d_o):
                b.free()
        return out

    def query_batch_device(self, n: int, d_bounds: "DeviceBuffer", d_visible: "DeviceBuffer",
                           d_clip: "DeviceBuffer | None" = None) -> None:
        _check(lib().rz_query_batch_device(self._ctx, n, d_bounds.ptr, d_visible.ptr,
                                           d_clip.ptr if d_clip is not None else None), "query_batch_device")

    def query_batch(self, bounds: np.ndarray):
        b = np.ascontiguousarray(np.asarray(bounds, dtype=np.float32).reshape(-1, 8))
        n = b.shape[0]
        vis = np.zeros(max(n, 1), np.uint8)
        clip = np.zeros(max(n, 1), np.uint8)
        _check(lib().rz_query_batch(self._ctx, n, _ptr(b), _ptr(vis), _ptr(clip)), "query_batch")
        return vis[:n].astype(bool), clip[:n].astype(bool)

    def raw(self):
        depth = np.zeros(self.blocks_x * self.blocks_y * 64, np.uint16)
        hiz = np.zeros(self.blocks_x * self.blocks_y + 8, np.uint16)
        _check(lib().rz_readback_raw(self._ctx, _ptr(depth), _ptr(hiz)), "readback_raw")
        return depth, hiz

    def upload_raw(self, depth=None, hiz=None) -> None:
        d = None if depth is None else np.ascontiguousarray(depth, dtype=np.uint16)
        h = None if hiz is None else np.ascontiguousarray(hiz, dtype=np.uint16)
        _check(lib().rz_upload_raw(self._ctx, _ptr(d) if d is not None else None,
                                   _ptr(h) if h is not None else None), "upload_raw")

    def tables(self):
        cov = np.zeros(4096, np.uint64)
        rcp = np.zeros(4096, np.uint32)
        _check(lib().rz_get_tables(self._ctx, _ptr(cov), _ptr(rcp)), "rz_get_tables")
        return cov, rcp

    def mvp_state(self):
        pre = np.zeros(16, np.float32)
        raw = np.zeros(16, np.float32)
        _check(lib().rz_get_mvp(self._ctx, _ptr(pre), _ptr(raw)), "rz_get_mvp")
        return pre, raw

    def sync(self) -> None:
        _check(lib().rz_sync(self._ctx), "rz_sync")

    def timer_start(self) -> None:
        _check(lib().rz_timer_start(self._ctx), "rz_timer_start")

    def profile(self, enable: bool) -> None:
        _check(lib().rz_profile_enable(self._ctx, 1 if enable else 0), "rz_profile_enable")

    def profile_read(self) -> dict:
        ms = np.zeros(4, np.float64)
        counts = np.zeros(2, np.uint32)
        _check(lib().rz_profile_read(self._ctx, _ptr(ms), _ptr(counts)), "rz_profile_read")
        return {"prep_ms": ms[0], "setup_ms": ms[1], "frame_ms": ms[2], "query_ms": ms[3],
                "draws": int(counts[0]), "query_batches": int(counts[1])}

    def frame_stats(self, enable: bool) -> None:
        _check(lib().rz_frame_stats_enable(self._ctx, 1 if enable else 0), "rz_frame_stats_enable")

    def frame_stats_read(self) -> np.ndarray:
        """Per-tile counters of the last draw, shape (tiles, 22): total ticks, wait
        ticks, waits, query ticks, queries, raster ticks, quads, items, scan
        ticks, scan passes, load ticks, item-loop ticks, draw ticks, visited,
        rejected and written (quad, block) pairs, fold-staging ticks,
        block-loop ticks, chunks, fold additions, early query partials tried
        and answered."""
        n = _U32(0)
        _check(lib().rz_frame_stats_read(self._ctx, None, ctypes.byref(n)), "rz_frame_stats_read")
        out = np.zeros((max(n.value, 1), 22), np.uint64)
        _check(lib().rz_frame_stats_read(self._ctx, _ptr(out), ctypes.byref(n)), "rz_frame_stats_read")
        return out[: n.value]

    def timer_stop(self) -> float:
        ms = ctypes.c_float(0.0)
        _check(lib().rz_timer_stop(self._ctx, ctypes.byref(ms)), "rz_timer_stop")
        return float(ms.value)


class Views:
    """Batched multi-view launches (rz_views_*): per view, clear, classify each
    occluder with the depth-independent part of queryVisibility and rasterize
    it, then query the view's boxes; all views in one step.  `owner` is a
    Rasterizer whose context uploads the occluders (same device)."""

    def __init__(self, owner: "Rasterizer", sizes):
        self._owner = owner
        w = np.ascontiguousarray([int(a) for a, _ in sizes], np.uint32)
        h = np.ascontiguousarray([int(b) for _, b in sizes], np.uint32)
        self._v = ctypes.c_void_p()
        _check(lib().rz_views_create(len(sizes), _ptr(w), _ptr(h), ctypes.byref(self._v)), "rz_views_create")
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self._occ = [[] for _ in sizes]
        self._nq = [0] * len(sizes)
        _live[1].add(self)

    def close(self) -> None:
        if self._v and _lib is not None:
            _lib.rz_views_destroy(self._v)
        self._v = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mvp(self, view: int, matrix) -> None:
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.float32).reshape(16))
        _check(lib().rz_views_set_mvp(self._v, view, _ptr(m)), "rz_views_set_mvp")

    def set_occluders(self, view: int, occluders) -> None:
        self._occ[view] = list(occluders)  # keeps the device copies alive
        handles = self._owner.handle_array(occluders)
        _check(lib().rz_views_set_occluders(self._v, view, len(occluders), ctypes.cast(handles, ctypes.c_void_p)),
               "rz_views_set_occluders")

    def set_queries(self, view: int, bounds) -> None:
        b = np.ascontiguousarray(np.asarray(bounds, dtype=np.float32).reshape(-1, 8))
        _check(lib().rz_views_set_queries(self._v, view, b.shape[0], _ptr(b)), "rz_views_set_queries")
        self._nq[view] = b.shape[0]

    def run(self, graph: bool = True) -> None:
        _check(lib().rz_views_run(self._v, 1 if graph else 0), "rz_views_run")

    def sync(self) -> float:
        ms = ctypes.c_float(0.0)
        _check(lib().rz_views_sync(self._v, ctypes.byref(ms)), "rz_views_sync")
        return float(ms.value)

    def raw(self, view: int):
        w, h = self.sizes[view]
        nb = (w // 8) * (h // 8)
        depth = np.zeros(nb * 64, np.uint16)
        hiz = np.zeros(nb + 8, np.uint16)
        _check(lib().rz_views_read_raw(self._v, view, _ptr(depth), _ptr(hiz)), "rz_views_read_raw")
        return depth, hiz

    def queries(self, view: int):
        n = self._nq[view]
        vis = np.zeros(max(n, 1), np.uint8)
        clip = np.zeros(max(n, 1), np.uint8)
        _check(lib().rz_views_read_queries(self._v, view, _ptr(vis), _ptr(clip)), "rz_views_read_queries")
        return vis[:n].astype(bool), clip[:n].astype(bool)

    def status(self, view: int) -> np.ndarray:
        n = len(self._occ[view])
        out = np.zeros(max(n, 1), np.uint8)
        _check(lib().rz_views_read_status(self._v, view, _ptr(out)), "rz_views_read_status")
        return out[:n]

    def enable_labels(self, enable: bool = True) -> None:
        _check(lib().rz_views_enable_labels(self._v, 1 if enable else 0), "rz_views_enable_labels")

    def n_batches(self, view: int) -> int:
        return sum(int(o.packets.size) // 32 for o in self._occ[view])

    def set_batch_order(self, view: int, order=None) -> None:
        o = np.ascontiguousarray(np.zeros(0) if order is None else order, np.uint32)
        _check(lib().rz_views_set_batch_order(self._v, view, o.size, _ptr(o) if o.size else None),
               "rz_views_set_batch_order")

    def labels(self, view: int):
        n = self.n_batches(view)
        lab = np.zeros(max(n, 1), np.uint32)
        feat = np.zeros((max(n, 1), 16), np.float32)
        _check(lib().rz_views_read_labels(self._v, view, _ptr(lab), _ptr(feat)), "rz_views_read_labels")
        return lab[:n], feat[:n]

    def set_model(self, mlp: "Mlp | None", by_occluder: bool = False) -> None:
        """Learned order: batches by descending score, or with by_occluder
        whole occluders by their best batch's score.  Batches of culled
        occluders score -inf (occluders: -FLT_MAX) and come last."""
        self._model = mlp
        self._by_occluder = by_occluder and mlp is not None
        _check(lib().rz_views_set_model(self._v, mlp._m if mlp is not None else None, 1 if by_occluder else 0),
               "rz_views_set_model")

    def enable_stats(self, enable: bool = True) -> None:
        _check(lib().rz_views_enable_stats(self._v, 1 if enable else 0), "rz_views_enable_stats")

    def stats(self) -> dict:
        """Block-loop counters of the last step over all views."""
        a = np.zeros(3, np.uint64)
        _check(lib().rz_views_read_stats(self._v, _ptr(a)), "rz_views_read_stats")
        return {"visited": int(a[0]), "rejected": int(a[1]), "written": int(a[2])}

    def tile_stats(self) -> np.ndarray:
        """Per-tile counters of the last step summed over all units (22 words,
        the layout of Rasterizer.frame_stats_read; word 0: work-group
        lifetimes)."""
        a = np.zeros(22, np.uint64)
        _check(lib().rz_views_read_tile_stats(self._v, _ptr(a)), "rz_views_read_tile_stats")
        return a

    def set_eye(self, view: int, eye) -> None:
        """Front-to-back order from `eye` computed on the device every step
        (Main.cpp:184-190, ties in submission order); None restores
        submission order."""
        if eye is None:
            _check(lib().rz_views_set_eye(self._v, view, None), "rz_views_set_eye")
            getattr(self, "_eyes", set()).discard(view)
            return
        e = np.ascontiguousarray(np.asarray(eye, np.float32)[:3])
        _check(lib().rz_views_set_eye(self._v, view, _ptr(e)), "rz_views_set_eye")
        if not hasattr(self, "_eyes"):
            self._eyes = set()
        self._eyes.add(view)

    def order(self, view: int) -> np.ndarray:
        by_occ = getattr(self, "_by_occluder", False) or bool(getattr(self, "_eyes", set()))
        n = len(self._occ[view]) if by_occ else self.n_batches(view)
        out = np.zeros(max(n, 1), np.uint32)
        _check(lib().rz_views_read_order(self._v, view, _ptr(out)), "rz_views_read_order")
        return out[:n]

    STATE_DTYPE = np.dtype([("T", "<f4", (4, 4)), ("c0", "<f4"), ("c1", "<f4"), ("status", "<u4"),
                            ("q", "<u4", (5,)), ("expected", "<u4"), ("fp", "<u4", (4,)), ("pad", "<u4")])
    SETUP_DTYPE = np.dtype([("nx", "<f4", (4,)), ("ny", "<f4", (4,)), ("off", "<f4", (4,)), ("dp0", "<f4"),
                            ("dDx", "<f4"), ("dDy", "<f4"), ("fb", "<u4"), ("range", "<u4"), ("modeZ", "<u4"),
                            ("slope01", "<u4"), ("slope23", "<u4")])

    def items(self, view: int):
        """The last step's per-occluder prologue state and per-quad setup
        records of a view (rz_views_read_items), as structured arrays."""
        occ = self._occ[view]
        st = np.zeros(max(len(occ), 1), self.STATE_DTYPE)
        nq = sum(int(o.packets.size) // 4 for o in occ)
        rec = np.zeros(max(nq, 1), self.SETUP_DTYPE)
        _check(lib().rz_views_read_items(self._v, view, _ptr(st), _ptr(rec)), "rz_views_read_items")
        return st[: len(occ)], rec[:nq]

    def profile(self, enable: bool = True) -> None:
        _check(lib().rz_views_profile_enable(self._v, 1 if enable else 0), "rz_views_profile_enable")

    def profile_read(self) -> dict:
        """Milliseconds of the last step's phases: setup (clear, prep, setup
        and any ordering work), raster (the persistent launch) and queries."""
        a = np.zeros(3, np.float32)
        _check(lib().rz_views_profile_read(self._v, _ptr(a)), "rz_views_profile_read")
        return {"setup_ms": float(a[0]), "raster_ms": float(a[1]), "query_ms": float(a[2])}

    def set_query_gated(self, gated: bool = True) -> None:
        """Every view runs the reference's query-gated draw loop (as rz_draw_frame)."""
        _check(lib().rz_views_set_query_gated(self._v, 1 if gated else 0), "rz_views_set_query_gated")

    def set_wait_timeout(self, ticks: int) -> None:
        """Bound of each cross-tile wait of query-gated steps in 100 MHz ticks (0: the default 2 s)."""
        _check(lib().rz_views_set_wait_timeout(self._v, int(ticks)), "rz_views_set_wait_timeout")

    def config(self) -> dict:
        c = ViewsConfig()
        _check(lib().rz_views_get_config(self._v, ctypes.byref(c)), "rz_views_get_config")
        return {"tileW": c.tileW, "waves": c.waves, "groupsPerCU": c.groupsPerCU}

    def set_config(self, waves: int, groups_per_cu: int = 0, tile_w: int = 8) -> None:
        _check(lib().rz_views_set_config(self._v, tile_w, waves, groups_per_cu), "rz_views_set_config")

    def autotune(self, reps: int = 5, cache: "str | os.PathLike | None" = None) -> dict:
        """rz_views_tune: times the current step under every configuration
        (identical bits required) and keeps the fastest."""
        cands = (ViewsConfig * 12)()
        k = ctypes.c_uint32(0)
        _check(lib().rz_views_tune(self._v, reps, os.fsencode(cache) if cache is not None else None, cands,
                                   ctypes.byref(k)), "rz_views_tune")
        return {"chosen": self.config(), "candidates": [_cfg_dict(cands[i]) for i in range(k.value)]}

    def load_config(self, cache) -> bool:
        found = ctypes.c_int(0)
        _check(lib().rz_views_load_config(self._v, os.fsencode(cache), ctypes.byref(found)), "rz_views_load_config")
        return bool(found.value)

    def info(self) -> dict:
        a = np.zeros(8, np.uint32)
        _check(lib().rz_views_get_info(self._v, _ptr(a)), "rz_views_get_info")
        keys = ["views", "items", "quads", "units", "queries", "groups", "graph", "dirty"]
        return dict(zip(keys, (int(x) for x in a)))


class Mlp:
    """The batch-scoring MLP (rz_mlp_*): 16 -> 64 -> 64 -> 1, GELU, BF16 MFMA."""

    N_PARAMS = 5313

    def __init__(self, seed: int = 1):
        self._m = ctypes.c_void_p()
        _check(lib().rz_mlp_create(seed, ctypes.byref(self._m)), "rz_mlp_create")
        self.uses_labels = True
        _live[1].add(self)

    def close(self) -> None:
        if self._m and _lib is not None:
            _lib.rz_mlp_destroy(self._m)
        self._m = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        w = np.zeros(self.N_PARAMS, np.float32)
        n = np.zeros(32, np.float32)
        _check(lib().rz_mlp_get_params(self._m, _ptr(w), _ptr(n)), "rz_mlp_get_params")
        return w, n

    def set_params(self, params=None, norm=None) -> None:
        w = None if params is None else np.ascontiguousarray(params, np.float32).reshape(self.N_PARAMS)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32).reshape(32)
        _check(lib().rz_mlp_set_params(self._m, _ptr(w) if w is not None else None,
                                       _ptr(n) if n is not None else None), "rz_mlp_set_params")

    def set_uses_labels(self, uses: bool) -> None:
        """False for a model trained with feature 10 (the previous step's
        label) held at 0: ordered steps then skip the label tracking."""
        self.uses_labels = bool(uses)
        _check(lib().rz_mlp_set_uses_labels(self._m, 1 if uses else 0), "rz_mlp_set_uses_labels")

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        out = np.zeros(max(x.shape[0], 1), np.float32)
        _check(lib().rz_mlp_forward(self._m, x.shape[0], _ptr(x), _ptr(out)), "rz_mlp_forward")
        return out[: x.shape[0]]

    def gradients(self, x, pairs):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().rz_mlp_gradients(self._m, x.shape[0], _ptr(x), pr.shape[0], _ptr(pr), _ptr(g), _ptr(loss)),
               "rz_mlp_gradients")
        return g, float(loss[0])

    def train_step(self, x, pairs, lr: float) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 16)
        pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        loss = np.zeros(1, np.float32)
        _check(lib().rz_mlp_train_step(self._m, x.shape[0], _ptr(x), pr.shape[0], _ptr(pr), ctypes.c_float(lr),
                                       _ptr(loss)), "rz_mlp_train_step")
        return float(loss[0])


class QuadDecomposition:
    """QuadDecomposition (QuadDecomposition.h:5-11) under the reference's name."""

    @staticmethod
    def decompose(indices, vertices) -> np.ndarray:
        """Triangle indices + vertices (n x 4 or n x 3 floats) -> quad indices
        (4 per quad), identical to the reference's QuadDecomposition::decompose."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
        v = np.asarray(vertices, dtype=np.float32)
        if v.ndim == 2 and v.shape[1] == 3:
            v = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
        v = np.ascontiguousarray(v.reshape(-1, 4))
        out = np.zeros(max(4 * (idx.size // 3), 1), np.uint32)
        n = ctypes.c_size_t(0)
        _check(lib().rz_quad_decompose(_ptr(idx), idx.size, _ptr(v), v.shape[0], _ptr(out), ctypes.byref(n)),
               "QuadDecomposition.decompose")
        return out[:n.value].copy()


class SurfaceAreaHeuristic:
    """SurfaceAreaHeuristic (SurfaceAreaHeuristic.h:5-11) with its static
    generateBatches computed on the GPU (rz_sah_batches)."""

    _rast = None

    @staticmethod
    def generateBatches(aabbs, targetSize: int, splitGranularity: int) -> list:
        """aabbs: (n, 8) or (n, 2, 4) boxes (min xyzw, max xyzw), as
        std::vector<Aabb> holds them; returns lists of quad indices in the
        reference's batch order (SurfaceAreaHeuristic.cpp:96-104)."""
        b = np.asarray(aabbs, np.float32).reshape(-1, 8)
        if SurfaceAreaHeuristic._rast is None:
            SurfaceAreaHeuristic._rast = Rasterizer(8, 8)
        out = SurfaceAreaHeuristic._rast.sah_batches_device(b[:, 0:4], b[:, 4:8], targetSize, splitGranularity)
        return [x.tolist() for x in out]


__all__ = ["Rasterizer", "Occluder", "SurfaceAreaHeuristic", "DeviceBuffer", "Views", "Mlp", "RasterizerError", "lib",
           "LIB_PATH"]

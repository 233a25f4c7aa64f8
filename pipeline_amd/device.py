"""Device buffers, streams, events and graphs over the C ABI, plus the
SoA upload of scenes (matrices in row planes, boxes in two arrays)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check


class DeviceArray:
    """A device allocation with a numpy dtype and shape (device memory via dpc_malloc)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(_lib.load().dpc_malloc(C.byref(self.ptr), self.nbytes))

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def offset(self, elements: int) -> int:
        return self.addr + int(elements) * self.dtype.itemsize

    @classmethod
    def from_host(cls, arr: np.ndarray, stream=None) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr, stream)
        return d

    def upload(self, arr: np.ndarray, stream=None) -> None:
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.nbytes > self.nbytes:
            raise ValueError("upload larger than the allocation")
        check(_lib.load().dpc_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes, stream))
        if stream is None:
            check(_lib.load().dpc_stream_sync(None))

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(_lib.load().dpc_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        check(_lib.load().dpc_stream_sync(stream))
        return out

    def fill_u32(self, value: int, stream=None) -> None:
        check(_lib.load().dpc_fill_u32(self.ptr, int(value) & 0xFFFFFFFF, self.nbytes // 4, stream))

    def free(self) -> None:
        if self.ptr.value:
            check(_lib.load().dpc_free(self.ptr))
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            if self.ptr.value and _lib._lib is not None:
                _lib._lib.dpc_free(self.ptr)
        except Exception:
            pass


def workspace(nbytes: int, stream=None) -> DeviceArray:
    """A device workspace for dpc_result_diff / dpc_bbox: zero before its first
    use (the calls keep their state in it and leave it consistent)."""
    ws = DeviceArray(max((int(nbytes) + 7) // 8, 1), np.uint64)
    ws.fill_u32(0, stream)
    if stream is None:
        check(_lib.load().dpc_stream_sync(None))
    return ws


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        check(_lib.load().dpc_stream_create(C.byref(self.handle)))

    def sync(self) -> None:
        check(_lib.load().dpc_stream_sync(self.handle))

    def destroy(self) -> None:
        if self.handle.value:
            check(_lib.load().dpc_stream_destroy(self.handle))
            self.handle = C.c_void_p()


class Event:
    def __init__(self):
        self.handle = C.c_void_p()
        check(_lib.load().dpc_event_create(C.byref(self.handle)))

    def record(self, stream=None) -> None:
        check(_lib.load().dpc_event_record(self.handle, stream))

    def sync(self) -> None:
        check(_lib.load().dpc_event_sync(self.handle))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(_lib.load().dpc_event_elapsed_ms(self.handle, later.handle, C.byref(ms)))
        return float(ms.value)

    def destroy(self) -> None:
        if self.handle.value:
            check(_lib.load().dpc_event_destroy(self.handle))
            self.handle = C.c_void_p()


class Graph:
    """hipGraph captured from work issued on a stream."""

    def __init__(self, stream: Stream, issue):
        lib = _lib.load()
        check(lib.dpc_graph_begin(stream.handle))
        try:
            issue()
        finally:
            self.handle = C.c_void_p()
            rc = lib.dpc_graph_end(stream.handle, C.byref(self.handle))
        check(rc)

    def launch(self, stream: Stream) -> None:
        check(_lib.load().dpc_graph_launch(self.handle, stream.handle))

    def destroy(self) -> None:
        if self.handle.value:
            check(_lib.load().dpc_graph_destroy(self.handle))
            self.handle = C.c_void_p()


def device_info(device: int = 0) -> dict:
    lib = _lib.load()
    name = C.create_string_buffer(256)
    cus, clk, l2 = C.c_int(), C.c_int(), C.c_int()
    mem = C.c_size_t()
    check(lib.dpc_device_info(device, name, 256, C.byref(cus), C.byref(clk), C.byref(mem), C.byref(l2)))
    return {"name": name.value.decode(), "cu_count": cus.value, "clock_khz": clk.value,
            "total_mem": mem.value, "l2_bytes": l2.value}


def mask_words(n: int) -> int:
    """64-bit words of a visibility mask of n objects."""
    return (int(n) + 63) // 64


@dataclass
class DeviceScene:
    """Objects and matrices resident on the device in the backend's SoA layout."""
    box_a: DeviceArray      # n x float4 (lower.xyz, extent.x)
    box_b: DeviceArray      # n x float2 (extent.y, extent.z)
    planes: DeviceArray     # 4 x ld x float4 row planes
    ld: int
    n_objects: int
    xform: DeviceArray | None = None  # n x uint32 matrix index, None = object i uses matrix i

    @property
    def xform_addr(self) -> int | None:
        return self.xform.addr if self.xform is not None else None

    def free(self) -> None:
        for a in (self.box_a, self.box_b, self.planes, self.xform):
            if a is not None:
                a.free()


def matrices_to_planes(m: np.ndarray, ld: int | None = None) -> np.ndarray:
    """(count, 4, 4) row-major matrices -> (4, ld, 4) row planes (host-side form of the device layout)."""
    m = np.asarray(m, np.float32).reshape(-1, 4, 4)
    ld = m.shape[0] if ld is None else ld
    planes = np.zeros((4, ld, 4), np.float32)
    planes[:, : m.shape[0], :] = np.transpose(m, (1, 0, 2))
    return planes


def upload_scene(lower: np.ndarray, upper: np.ndarray, matrices: np.ndarray,
                 xform: np.ndarray | None = None, stream=None) -> DeviceScene:
    """Upload boxes (lower/upper, n x 3) and matrices (count x 4 x 4, AoS as the
    reference stores them).  The extent subtraction and the AoS -> row-plane
    transpose run on the device (dpc_boxes_from_lower_upper, dpc_matrices_from_aos)."""
    lib = _lib.load()
    n = int(lower.shape[0])
    lu = np.ascontiguousarray(np.concatenate([lower, upper], axis=1), np.float32)
    m = np.ascontiguousarray(matrices, np.float32).reshape(-1, 16)
    ld = m.shape[0]
    d_lu = DeviceArray.from_host(lu, stream)
    box_a = DeviceArray((max(n, 1), 4), np.float32)
    box_b = DeviceArray((max(n, 1), 2), np.float32)
    check(lib.dpc_boxes_from_lower_upper(d_lu.ptr, n, box_a.ptr, box_b.ptr, stream))
    d_aos = DeviceArray.from_host(m, stream)
    planes = DeviceArray((4, max(ld, 1), 4), np.float32)
    check(lib.dpc_matrices_from_aos(d_aos.ptr, 64, ld, planes.ptr, ld, stream))
    check(lib.dpc_stream_sync(stream))
    d_lu.free()
    d_aos.free()
    dx = DeviceArray.from_host(np.ascontiguousarray(xform, np.uint32), stream) if xform is not None else None
    return DeviceScene(box_a, box_b, planes, ld, n, dx)


def vp_arg(vp: np.ndarray):
    v = np.ascontiguousarray(vp, np.float32).reshape(16)
    return (C.c_float * 16)(*v.tolist())


def cull(scene: DeviceScene, vp: np.ndarray, out: DeviceArray | None = None, stream=None) -> DeviceArray:
    """dpc_cull over the whole scene; returns the device mask (ceil(n/64) uint64 words)."""
    lib = _lib.load()
    if out is None:
        out = DeviceArray(max(mask_words(scene.n_objects), 1), np.uint64)
    check(lib.dpc_cull(scene.box_a.ptr, scene.box_b.ptr, scene.xform_addr, scene.n_objects, scene.planes.ptr,
                       scene.ld, vp_arg(vp), out.ptr, stream))
    return out


def scene_desc(scene: DeviceScene) -> "_lib.SceneDesc":
    return _lib.SceneDesc(scene.box_a.addr, scene.box_b.addr, scene.xform_addr or None, scene.planes.addr,
                          scene.ld, scene.n_objects)


class BatchPlan:
    """dpc_batch_plan: groups = (scene, first_object, count), pairs = (frustum, group, mask word offset)."""

    def __init__(self, groups, pairs, n_frusta: int, cfg=None):
        lib = _lib.load()
        G, P = len(groups), len(pairs)
        g = (_lib.GroupRange * max(G, 1))(*[_lib.GroupRange(int(s), 0, int(a), int(n)) for s, a, n in groups])
        p = (_lib.CullPair * max(P, 1))(*[_lib.CullPair(int(f), int(gi), int(o)) for f, gi, o in pairs])
        self.handle = C.c_void_p()
        check(lib.dpc_batch_plan_create(g, G, p, P, int(n_frusta), C.byref(cfg) if cfg is not None else None,
                                        C.byref(self.handle)))
        mw, tests, items = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib.dpc_batch_plan_info(self.handle, C.byref(mw), C.byref(tests), C.byref(items)))
        self.mask_words, self.tests, self.items = mw.value, tests.value, items.value

    def bind(self, scenes) -> None:
        table = (_lib.SceneDesc * max(len(scenes), 1))(*[scene_desc(s) for s in scenes])
        check(_lib.load().dpc_batch_plan_bind(self.handle, table, len(scenes)))

    def cull(self, d_vps: DeviceArray, d_masks: DeviceArray, stream=None) -> None:
        check(_lib.load().dpc_cull_batch(self.handle, d_vps.ptr, d_masks.ptr, d_masks.nbytes // 8, stream))

    def cull_per_view(self, h_vps: np.ndarray, d_masks: DeviceArray, stream=None) -> None:
        self._h_vps = np.ascontiguousarray(h_vps, np.float32).reshape(-1)
        check(_lib.load().dpc_cull_per_view(self.handle, self._h_vps.ctypes.data_as(C.POINTER(C.c_float)),
                                            d_masks.ptr, d_masks.nbytes // 8, stream))

    def destroy(self) -> None:
        if self.handle.value:
            _lib.load().dpc_batch_plan_destroy(self.handle)
            self.handle = C.c_void_p()


def batched_plan(bs, cfg=None, scene_index: int = 0) -> BatchPlan:
    """The plan of a scenes.BatchedScene whose objects form one device scene."""
    groups = [(scene_index, a, n) for a, n in zip(bs.group_first, bs.group_count)]
    pairs = list(zip(bs.pair_frustum, bs.pair_group, bs.pair_offset))
    return BatchPlan(groups, pairs, bs.vps.shape[0], cfg)

"""The serving loop over the C ABI (dpc_server_* in include/dpcull_hip.h).

A `Server` keeps scenes resident in device memory and culls requests of
different scenes together in one launch.  A request applies a matrix delta to
its scene and then tests the whole scene against 1..64 frusta.  The scheduler
is native (pipeline_amd/csrc/host/Serving.cpp): it runs a thread of its own,
and submissions from Python return at once.

The reference has no serving loop.  It culls one scene per frame on the render
thread (dp/sg/renderer/rix/gl/src/SceneRendererImpl.cpp:333-356, which calls
CullingImpl::cull, dp/sg/xbar/culling/src/CullingImpl.cpp:126-144).  Each
request's masks are what that call would produce for its scene, with the
deltas of the scene's earlier requests already applied.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check


class PinnedArray:
    """Pinned host memory (dpc_host_alloc) viewed as a numpy array.  Copies
    between it and the device are DMA transfers."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        nbytes = max(int(np.prod(shape, dtype=np.int64)) * self.dtype.itemsize, 1)
        self._ptr = C.c_void_p()
        check(_lib.load().dpc_host_alloc(C.byref(self._ptr), nbytes))
        buf = (C.c_char * nbytes).from_address(self._ptr.value)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=nbytes)[: int(np.prod(shape)) * self.dtype.itemsize] \
            .view(self.dtype).reshape(shape)

    def free(self) -> None:
        if self._ptr.value:
            self.array = None
            check(_lib.load().dpc_host_free(self._ptr))
            self._ptr = C.c_void_p()


@dataclass
class RequestResult:
    request_id: int
    scene: int
    status: int
    submit_us: float
    admit_us: float
    done_us: float
    batch: int
    batch_requests: int
    mask_words: int

    @property
    def latency_us(self) -> float:
        return self.done_us - self.submit_us


class Server:
    """dpc_server: resident scenes, admission against a device-memory budget,
    continuous batching of cull requests across scenes."""

    def __init__(self, max_scenes: int = 16, max_batch_requests: int = 32, streams: int = 4,
                 objects_per_lane: int = 1, waves_per_cu: int = 0, hbm_budget_bytes: int = 16 << 30,
                 batches_per_stream: int = 2, direct_delta_reads: bool = False):
        self.lib = _lib.load()
        cfg = _lib.ServerConfig(max_scenes, max_batch_requests, streams, objects_per_lane, waves_per_cu,
                                batches_per_stream, int(direct_delta_reads), hbm_budget_bytes)
        self.handle = C.c_void_p()
        check(self.lib.dpc_server_create(C.byref(cfg), C.byref(self.handle)))
        self.scene_objects: dict[int, int] = {}
        self.scene_matrices: dict[int, int] = {}
        self._borrowed: dict[int, tuple] = {}  # request id -> arrays the server reads until the request is done
        self._scene_of: dict[int, int] = {}

    # ---- scenes ------------------------------------------------------------
    def add_scene(self, lower: np.ndarray, upper: np.ndarray, matrices: np.ndarray,
                  xform: np.ndarray | None = None) -> int:
        """Uploads a scene (boxes as lower/upper n x 3, matrices k x 4 x 4 as the
        reference stores them, object i uses matrix xform[i] or i)."""
        lu = np.ascontiguousarray(np.concatenate([lower, upper], axis=1), np.float32)
        m = np.ascontiguousarray(matrices, np.float32).reshape(-1, 16)
        x = None if xform is None else np.ascontiguousarray(xform, np.uint32)
        sid = C.c_uint32()
        check(self.lib.dpc_server_add_scene(self.handle, lu.ctypes.data, lu.shape[0], m.ctypes.data, m.shape[0],
                                            None if x is None else x.ctypes.data, C.byref(sid)))
        self.scene_objects[sid.value] = int(lu.shape[0])
        self.scene_matrices[sid.value] = int(m.shape[0])
        return sid.value

    def remove_scene(self, scene: int) -> None:
        check(self.lib.dpc_server_remove_scene(self.handle, scene))
        del self.scene_objects[scene]
        del self.scene_matrices[scene]

    def read_matrices(self, scene: int) -> np.ndarray:
        """The scene's current matrices (k x 16 float32, row-major), every delta of
        its completed requests applied; the scene must have none in flight."""
        out = np.empty((self.scene_matrices[scene], 16), np.float32)
        check(self.lib.dpc_server_read_matrices(self.handle, scene, out.ctypes.data, out.shape[0]))
        return out

    def mask_words(self, scene: int, n_frusta: int) -> int:
        return n_frusta * ((self.scene_objects[scene] + 63) // 64)

    # ---- requests ----------------------------------------------------------
    def submit(self, scene: int, vps: np.ndarray, masks_out: np.ndarray,
               delta_idx: np.ndarray | None = None, delta_mats: np.ndarray | None = None) -> int:
        """Queues a request; `masks_out` (uint64, at least mask_words(scene, n_frusta)
        words, ideally a PinnedArray) and the delta arrays are kept alive here until
        wait() returns for the request."""
        v = np.ascontiguousarray(vps, np.float32).reshape(-1, 16)
        if masks_out.dtype != np.uint64 or not masks_out.flags.c_contiguous:
            raise ValueError("masks_out must be a contiguous uint64 array")
        n_delta = 0 if delta_idx is None else int(delta_idx.shape[0])
        di = dm = None
        if n_delta:
            di = np.ascontiguousarray(delta_idx, np.uint32)
            dm = np.ascontiguousarray(delta_mats, np.float32).reshape(-1, 16)
            if dm.shape[0] != n_delta:
                raise ValueError("delta_idx and delta_mats differ in length")
        rid = C.c_uint64()
        check(self.lib.dpc_server_submit(self.handle, scene, v.ctypes.data, v.shape[0],
                                         None if di is None else di.ctypes.data,
                                         None if dm is None else dm.ctypes.data, n_delta, masks_out.ctypes.data,
                                         masks_out.size, C.byref(rid)))
        self._borrowed[rid.value] = (v, di, dm, masks_out)
        self._scene_of[rid.value] = scene
        return rid.value

    def wait(self, request_id: int, timeout_ms: float = -1.0) -> RequestResult:
        info = _lib.RequestInfo()  # zeroed; filled only once the request is done
        rc = self.lib.dpc_server_wait(self.handle, request_id, timeout_ms, C.byref(info))
        if rc != _lib.DPC_OK and info.done_us == 0.0:
            check(rc)  # timed out (or unknown id): a pending request keeps its arrays
        self._borrowed.pop(request_id, None)
        scene = self._scene_of.pop(request_id, -1)
        res = RequestResult(request_id, scene, int(info.status), info.submit_us, info.admit_us, info.done_us,
                            int(info.batch), int(info.batch_requests), int(info.mask_words))
        check(rc)
        return res

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(self.lib.dpc_server_stats_get(self.handle, C.byref(st)))
        out = {}
        for name, _ in _lib.ServerStats._fields_:
            v = getattr(st, name)
            out[name] = list(v) if isinstance(v, C.Array) else v
        return out

    def now_us(self) -> float:
        return float(self.lib.dpc_server_now_us(self.handle))

    def close(self) -> None:
        if self.handle.value:
            check(self.lib.dpc_server_destroy(self.handle))
            self.handle = C.c_void_p()
            self._borrowed.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

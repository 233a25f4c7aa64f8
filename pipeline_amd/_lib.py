"""ctypes binding of libdpcull_hip.so (the C ABI declared in include/dpcull_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libdpcull_hip.so"

DPC_OK = 0
DPC_ERR_INVALID_ARGUMENT = 1
DPC_ERR_HIP = 2
DPC_ERR_OUT_OF_MEMORY = 3
DPC_ERR_OVER_BUDGET = 4
DPC_ERR_NOT_FOUND = 5
DPC_ERR_STATE = 6
DPC_ERR_INTERNAL = 7


class DpcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dpcull error {code}: {msg}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [("box_a", C.c_void_p), ("box_b", C.c_void_p), ("xform", C.c_void_p), ("planes", C.c_void_p),
                ("ld", C.c_uint64), ("n_objects", C.c_uint64)]


class GroupRange(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("reserved", C.c_uint32), ("first_object", C.c_uint64),
                ("count", C.c_uint64)]


class CullPair(C.Structure):
    _fields_ = [("frustum", C.c_uint32), ("group", C.c_uint32), ("mask_word_offset", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [("max_scenes", C.c_uint32), ("max_batch_requests", C.c_uint32), ("streams", C.c_uint32),
                ("objects_per_lane", C.c_uint32), ("waves_per_cu", C.c_uint32), ("batches_per_stream", C.c_uint32),
                ("direct_delta_reads", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_us", C.c_double), ("admit_us", C.c_double), ("done_us", C.c_double),
                ("batch", C.c_uint32), ("batch_requests", C.c_uint32), ("mask_words", C.c_uint64),
                ("status", C.c_int32), ("reserved", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("batches", C.c_uint64), ("tests", C.c_uint64),
                ("delta_matrices", C.c_uint64), ("scene_bytes", C.c_uint64), ("peak_inflight_bytes", C.c_uint64),
                ("busy_us", C.c_double), ("issue_us", C.c_double),
                ("issue_max_us", C.c_double), ("issue_part_max_us", C.c_double * 5),
                ("delta_direct", C.c_uint64), ("delta_packed", C.c_uint64)]


class CullConfig(C.Structure):
    _fields_ = [("waves_per_cu", C.c_uint32)]


class BatchConfig(C.Structure):
    _fields_ = [("objects_per_lane", C.c_uint32), ("block_threads", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("max_frusta_per_item", C.c_uint32)]


P = C.c_void_p
SZ = C.c_size_t
I = C.c_int
U32 = C.c_uint32
U64 = C.c_uint64
FP = C.POINTER(C.c_float)

# name -> (restype, argtypes)
_SIGNATURES = {
    "dpc_abi_version": (I, []),
    "dpc_last_error": (C.c_char_p, []),
    "dpc_debug_inject_error": (I, [C.c_int64, I]),
    "dpc_device_info": (I, [I, C.c_char_p, SZ, C.POINTER(I), C.POINTER(I), C.POINTER(SZ), C.POINTER(I)]),
    "dpc_set_device": (I, [I]),
    "dpc_malloc": (I, [C.POINTER(P), SZ]),
    "dpc_free": (I, [P]),
    "dpc_host_alloc": (I, [C.POINTER(P), SZ]),
    "dpc_host_free": (I, [P]),
    "dpc_memcpy_h2d": (I, [P, P, SZ, P]),
    "dpc_memcpy_d2h": (I, [P, P, SZ, P]),
    "dpc_memcpy_d2d": (I, [P, P, SZ, P]),
    "dpc_fill_u32": (I, [P, U32, SZ, P]),
    "dpc_stream_create": (I, [C.POINTER(P)]),
    "dpc_stream_destroy": (I, [P]),
    "dpc_stream_sync": (I, [P]),
    "dpc_device_sync": (I, []),
    "dpc_event_create": (I, [C.POINTER(P)]),
    "dpc_event_destroy": (I, [P]),
    "dpc_event_record": (I, [P, P]),
    "dpc_event_sync": (I, [P]),
    "dpc_event_elapsed_ms": (I, [P, P, C.POINTER(C.c_float)]),
    "dpc_graph_begin": (I, [P]),
    "dpc_graph_end": (I, [P, C.POINTER(P)]),
    "dpc_graph_launch": (I, [P, P]),
    "dpc_graph_destroy": (I, [P]),
    "dpc_matrices_from_aos": (I, [P, SZ, SZ, P, SZ, P]),
    "dpc_matrices_to_aos": (I, [P, SZ, P, SZ, P, SZ, P]),
    "dpc_matrices_scatter": (I, [P, SZ, P, P, SZ, P]),
    "dpc_boxes_from_lower_upper": (I, [P, SZ, P, P, P]),
    "dpc_gather_u32": (I, [P, SZ, P, SZ, P, P]),
    "dpc_cull": (I, [P, P, P, SZ, P, SZ, FP, P, P]),
    "dpc_cull_get_config": (I, [C.POINTER(CullConfig)]),
    "dpc_cull_set_config": (I, [C.POINTER(CullConfig)]),
    "dpc_batch_plan_create": (I, [C.POINTER(GroupRange), SZ, C.POINTER(CullPair), SZ, U32,
                                  C.POINTER(BatchConfig), C.POINTER(P)]),
    "dpc_batch_plan_destroy": (I, [P]),
    "dpc_batch_plan_info": (I, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)]),
    "dpc_batch_plan_bind": (I, [P, C.POINTER(SceneDesc), SZ]),
    "dpc_cull_batch": (I, [P, P, P, SZ, P]),
    "dpc_cull_per_view": (I, [P, FP, P, SZ, P]),
    "dpc_tree_propagate": (I, [P, P, SZ, P, C.POINTER(U64), SZ, P, P, P]),
    "dpc_bytes_to_bits_indexed": (I, [P, P, SZ, P, SZ, P]),
    "dpc_result_diff_workspace": (SZ, [SZ]),
    "dpc_result_diff": (I, [P, P, SZ, P, P, I, P, P]),
    "dpc_result_diff_noalloc": (I, [P, P, SZ, P, P, I, P]),
    "dpc_probe_valu_rate": (I, [I, C.c_uint32, C.POINTER(C.c_double), P]),
    "dpc_cull_diff_workspace": (SZ, [SZ]),
    "dpc_cull_diff": (I, [P, P, P, SZ, P, SZ, FP, P, P, P, P, I, P, P]),
    "dpc_bbox_workspace": (SZ, [SZ]),
    "dpc_bbox": (I, [P, P, P, SZ, P, SZ, P, P, P]),
    "dpc_inverse_transpose": (I, [P, SZ, P, SZ, P, SZ, P, P]),
    "dpc_tree_plan_create": (I, [P, SZ, P, SZ, U32, C.POINTER(P)]),
    "dpc_tree_plan_destroy": (I, [P]),
    "dpc_tree_plan_info": (I, [P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U64)]),
    "dpc_tree_propagate_plan": (I, [P, P, P, P, P, P]),
    "dpc_server_create": (I, [C.POINTER(ServerConfig), C.POINTER(P)]),
    "dpc_server_destroy": (I, [P]),
    "dpc_server_add_scene": (I, [P, P, U64, P, U64, P, C.POINTER(U32)]),
    "dpc_server_remove_scene": (I, [P, U32]),
    "dpc_server_submit": (I, [P, U32, P, U32, P, P, U32, P, U64, C.POINTER(U64)]),
    "dpc_server_wait": (I, [P, U64, C.c_double, C.POINTER(RequestInfo)]),
    "dpc_server_stats_get": (I, [P, C.POINTER(ServerStats)]),
    "dpc_server_now_us": (C.c_double, [P]),
    "dpc_server_read_matrices": (I, [P, U32, P, U64]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the backend library (once) and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`. "
            "There is no CPU fallback for the HIP backend.")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    if rc != DPC_OK:
        msg = load().dpc_last_error()
        raise DpcError(rc, msg.decode() if msg else "")

"""ctypes wrapper of liboracle.so (the C restatement in oracle.c) and of the
compiled-reference companion oracle/_ref/libdpref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libdpref.so"

_lib = None
_ref = None

P = C.c_void_p
SZ = C.c_size_t


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
        lib = C.CDLL(str(LIB))
        sig = {
            "orc_vec4_mul_mat44": [P, P, P],
            "orc_mat44_mul": [P, P, P],
            "orc_mat44_mul_f64": [P, P, P],
            "orc_obb_build": [P, P, P, P],
            "orc_is_visible_obb": [P, P],
            "orc_obb_corners_clip": [P, P, P],
            "orc_cull_group": [SZ, P, P, P, P, SZ, P, P],
            "orc_obb_build_group": [SZ, P, P, P, P, SZ, P],
            "orc_cull_obbs": [SZ, P, P, P],
            "orc_bbox_group": [SZ, P, P, P, P, SZ, P],
            "orc_tree_compute": [SZ, P, P, P, P, SZ, P, P, P],
            "orc_tree_compute_f64": [SZ, P, P, P, P],
            "orc_result_diff": [SZ, P, P, P],
            "orc_invert_transpose": [P, P],
            "orc_obb_build_group_mt": [SZ, P, P, P, P, SZ, P, C.c_int],
            "orc_cull_obbs_mt": [SZ, P, P, P, C.c_int],
        }
        for name, args in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int if name in ("orc_is_visible_obb", "orc_invert_transpose") else \
                (SZ if name == "orc_result_diff" else None)
        _lib = lib
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def load_ref():
    """The reference's own Tree/BitArray/math code (built only where /root/reference exists)."""
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise FileNotFoundError(f"{REF_LIB} not built (needs the reference sources)")
        ref = C.CDLL(str(REF_LIB))
        u32p = C.POINTER(C.c_uint32)
        sig = {
            "ref_invert_transpose": (C.c_int, [P, P]),
            "ref_vec4_mul_mat44": (None, [P, P, P]),
            "ref_mat44_mul": (None, [P, P, P]),
            "ref_scalar_mul_vec4": (None, [C.c_float, P, P]),
            "ref_vec4_add": (None, [P, P, P]),
            "ref_make_perspective": (None, [C.c_float] * 4 + [P]),
            "ref_make_frustum": (None, [C.c_float] * 6 + [P]),
            "ref_make_ortho": (None, [C.c_float] * 6 + [P]),
            "ref_make_lookat": (None, [P, P, P, P]),
            "ref_set_mat_quat": (None, [P, P, P, P]),
            "ref_bitarray_setbits": (None, [P, SZ, P]),
            "ref_bitarray_count_leading_zeroes": (SZ, [P, SZ]),
            "ref_bitarray_traverse": (SZ, [P, SZ, P]),
            "ref_tree_create": (P, []),
            "ref_tree_destroy": (None, [P]),
            "ref_tree_add": (C.c_int, [P, C.c_uint32, P, u32p]),
            "ref_tree_remove": (C.c_int, [P, C.c_uint32]),
            "ref_tree_update_local": (None, [P, C.c_uint32, P]),
            "ref_tree_compute": (None, [P, P]),
            "ref_tree_count": (SZ, [P]),
            "ref_tree_world": (None, [P, P]),
            "ref_tree_level_count": (SZ, [P]),
            "ref_tree_level_size": (SZ, [P, SZ]),
            "ref_tree_level_entries": (None, [P, SZ, P]),
            "ref_tree_node_level": (C.c_uint32, [P, C.c_uint32]),
            "ref_tree_is_free": (C.c_int, [P, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(ref, name)
            fn.restype = res
            fn.argtypes = args
        _ref = ref
    return _ref


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


# ---------------------------------------------------------------------------
# single operations
# ---------------------------------------------------------------------------
def vec4_mul_mat44(v, m) -> np.ndarray:
    v, m = _f32(v).reshape(4), _f32(m).reshape(16)
    out = np.empty(4, np.float32)
    load().orc_vec4_mul_mat44(_p(v), _p(m), _p(out))
    return out


def mat44_mul(a, b) -> np.ndarray:
    a, b = _f32(a).reshape(16), _f32(b).reshape(16)
    out = np.empty(16, np.float32)
    load().orc_mat44_mul(_p(a), _p(b), _p(out))
    return out.reshape(4, 4)


def obb_build(lower4, extent4, m) -> np.ndarray:
    l, e, mm = _f32(lower4).reshape(4), _f32(extent4).reshape(4), _f32(m).reshape(16)
    out = np.empty(16, np.float32)
    load().orc_obb_build(_p(l), _p(e), _p(mm), _p(out))
    return out.reshape(4, 4)


def is_visible_obb(vp, obb) -> bool:
    vp, obb = _f32(vp).reshape(16), _f32(obb).reshape(16)
    return bool(load().orc_is_visible_obb(_p(vp), _p(obb)))


def corners_clip(vp, obb) -> np.ndarray:
    vp, obb = _f32(vp).reshape(16), _f32(obb).reshape(16)
    out = np.empty(32, np.float32)
    load().orc_obb_corners_clip(_p(vp), _p(obb), _p(out))
    return out.reshape(8, 4)


# ---------------------------------------------------------------------------
# bulk
# ---------------------------------------------------------------------------
def mask_u32_words(n: int) -> int:
    return 2 * ((int(n) + 63) // 64)


def cull_group(lower4, extent4, xform, matrices, vp) -> np.ndarray:
    """cpu::ManagerImpl::cull on one group; returns the mask as uint64 words."""
    l4, e4 = _f32(lower4).reshape(-1, 4), _f32(extent4).reshape(-1, 4)
    n = l4.shape[0]
    x = np.ascontiguousarray(xform, np.uint64).reshape(n)
    m = _f32(matrices).reshape(-1, 16)
    vp = _f32(vp).reshape(16)
    mask = np.zeros(max(mask_u32_words(n), 2), np.uint32)
    load().orc_cull_group(n, _p(l4), _p(e4), _p(x), _p(m), 64, _p(vp), _p(mask))
    return mask.view(np.uint64)[: (n + 63) // 64]


def obb_build_group(lower4, extent4, xform, matrices) -> np.ndarray:
    """GroupCPU::updateOBBs: (n, 4, 4) OBBs (point, ex, ey, ez)."""
    l4, e4 = _f32(lower4).reshape(-1, 4), _f32(extent4).reshape(-1, 4)
    n = l4.shape[0]
    x = np.ascontiguousarray(xform, np.uint64).reshape(n)
    m = _f32(matrices).reshape(-1, 16)
    out = np.empty((max(n, 1), 16), np.float32)
    load().orc_obb_build_group(n, _p(l4), _p(e4), _p(x), _p(m), 64, _p(out))
    return out[:n].reshape(n, 4, 4)


def cull_obbs(obbs, vp) -> np.ndarray:
    """The cull loop over cached OBBs (ManagerImpl.cpp:509-515); uint64 mask words."""
    o = _f32(obbs).reshape(-1, 16)
    n = o.shape[0]
    vp = _f32(vp).reshape(16)
    mask = np.zeros(max(mask_u32_words(n), 2), np.uint32)
    load().orc_cull_obbs(n, _p(o), _p(vp), _p(mask))
    return mask.view(np.uint64)[: (n + 63) // 64]


def bbox_group(lower4, extent4, xform, matrices) -> np.ndarray:
    l4, e4 = _f32(lower4).reshape(-1, 4), _f32(extent4).reshape(-1, 4)
    n = l4.shape[0]
    x = np.ascontiguousarray(xform, np.uint64).reshape(n)
    m = _f32(matrices).reshape(-1, 16)
    out = np.empty(6, np.float32)
    load().orc_bbox_group(n, _p(l4), _p(e4), _p(x), _p(m), 64, _p(out))
    return out


def tree_compute(level_offsets, entries, local, world, dirty_local, dirty_world):
    """Tree::compute on index-addressed AoS arrays (modified in place).
    Returns the EventWorldMatricesChanged bytes."""
    lo = np.ascontiguousarray(level_offsets, np.uint64)
    en = np.ascontiguousarray(entries, np.uint32)
    n_index = local.shape[0]
    assert local.dtype == np.float32 and world.dtype == np.float32 and local.flags.c_contiguous and world.flags.c_contiguous
    assert dirty_local.dtype == np.uint8 and dirty_world.dtype == np.uint8
    event = np.zeros(n_index, np.uint8)
    load().orc_tree_compute(lo.shape[0] - 1, _p(lo), _p(en), _p(local), _p(world), n_index, _p(dirty_local),
                            _p(dirty_world), _p(event))
    return event


def tree_compute_f64(level_offsets, entries, local64) -> np.ndarray:
    lo = np.ascontiguousarray(level_offsets, np.uint64)
    en = np.ascontiguousarray(entries, np.uint32)
    local64 = np.ascontiguousarray(local64, np.float64)
    world = np.zeros_like(local64)
    world[0] = np.eye(4)
    load().orc_tree_compute_f64(lo.shape[0] - 1, _p(lo), _p(en), _p(local64), _p(world))
    return world


def result_diff(n, old_words, new_words) -> np.ndarray:
    o = np.ascontiguousarray(old_words, np.uint64).view(np.uint32)
    w = np.ascontiguousarray(new_words, np.uint64).view(np.uint32)
    out = np.empty(max(int(n), 1), np.uint32)
    c = load().orc_result_diff(int(n), _p(o), _p(w), _p(out))
    return out[:c].copy()


def invert_transpose(ms):
    """dp::math::invertTranspose over (n, 4, 4) matrices: (ok (n,) uint8, out (n, 4, 4)).
    Where ok is 0 (det == 0) the output keeps the value of `out`'s initial fill (NaN)."""
    m = _f32(ms).reshape(-1, 16)
    out = np.full(m.shape, np.nan, np.float32)
    ok = np.zeros(m.shape[0], np.uint8)
    fn = load().orc_invert_transpose
    for i in range(m.shape[0]):
        ok[i] = fn(_p(m[i]), _p(out[i]))
    return ok, out.reshape(-1, 4, 4)


def obb_build_group_mt(lower4, extent4, xform, matrices, threads: int) -> np.ndarray:
    """obb_build_group over `threads` host threads (same result)."""
    l4, e4 = _f32(lower4).reshape(-1, 4), _f32(extent4).reshape(-1, 4)
    n = l4.shape[0]
    x = np.ascontiguousarray(xform, np.uint64).reshape(n)
    m = _f32(matrices).reshape(-1, 16)
    out = np.empty((max(n, 1), 16), np.float32)
    load().orc_obb_build_group_mt(n, _p(l4), _p(e4), _p(x), _p(m), 64, _p(out), int(threads))
    return out[:n].reshape(n, 4, 4)


def cull_obbs_mt(obbs, vp, threads: int) -> np.ndarray:
    """cull_obbs over `threads` host threads (same result)."""
    o = _f32(obbs).reshape(-1, 16)
    n = o.shape[0]
    vp = _f32(vp).reshape(16)
    mask = np.zeros(max(mask_u32_words(n), 2), np.uint32)
    load().orc_cull_obbs_mt(n, _p(o), _p(vp), _p(mask), int(threads))
    return mask.view(np.uint64)[: (n + 63) // 64]

"""The C-ABI library loads (no GPU needed) and exports every function that
include/dpcull_hip.h declares; the Python binding declares the same set; bad
arguments are rejected on the host without a GPU."""
import re

from conftest import ROOT


def declared_functions():
    text = (ROOT / "include" / "dpcull_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dpc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dpc_cull" in names and "dpc_cull_batch" in names and "dpc_tree_propagate" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from pipeline_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared_functions()


def test_abi_version_without_gpu(lib):
    assert lib.dpc_abi_version() == 2


def test_argument_errors_are_reported_before_any_device_call(lib):
    """Bad arguments come back as DPC_ERR_INVALID_ARGUMENT with a message, from
    host-side checks that run before the device is touched (the addresses below
    are never dereferenced)."""
    inval = 1  # DPC_ERR_INVALID_ARGUMENT
    a = 0x10000  # 256-byte aligned
    # dpc_bbox: workspace not 8-byte aligned, box_b not 8-byte aligned, ld < n without indices
    assert lib.dpc_bbox(a, a, None, 10, a, 10, a, a + 4, None) == inval
    assert b"aligned" in lib.dpc_last_error()
    assert lib.dpc_bbox(a, a + 4, None, 10, a, 10, a, a, None) == inval
    assert lib.dpc_bbox(a, a, None, 10, a, 5, a, a, None) == inval
    assert b"ld >= n" in lib.dpc_last_error()
    # dpc_cull: no view-projection matrix
    assert lib.dpc_cull(a, a, None, 10, a, 10, None, a, None) == inval
    # dpc_result_diff: no count output
    assert lib.dpc_result_diff(a, a, 100, a, None, 0, a, None) == inval
    # dpc_result_diff_noalloc: more bits than its fixed scratch serves
    assert lib.dpc_result_diff_noalloc(a, a, (1 << 30) + 64, a, a, 0, None) == inval
    assert b"exceed" in lib.dpc_last_error()
    # dpc_cull_diff: no workspace, a misaligned one, the old mask as the new one
    import ctypes
    vp = (ctypes.c_float * 16)()
    assert lib.dpc_cull_diff(a, a, None, 100, a, 100, vp, a, a + 4096, a, a, 1, None, None) == inval
    assert b"workspace" in lib.dpc_last_error()
    assert lib.dpc_cull_diff(a, a, None, 100, a, 100, vp, a, a + 4096, a, a, 1, a + 8, None) == inval
    assert lib.dpc_cull_diff(a, a, None, 100, a, 100, vp, a, a, a, a, 1, a, None) == inval
    assert b"one buffer" in lib.dpc_last_error()
    assert lib.dpc_cull_diff_workspace(1 << 22) >= (1 << 16)
    # dpc_inverse_transpose: more matrices than the leading dimension holds
    assert lib.dpc_inverse_transpose(a, 5, None, 10, a, 10, None, None) == inval
    assert b"exceeds" in lib.dpc_last_error()
    # dpc_server_read_matrices: no server, no output array
    assert lib.dpc_server_read_matrices(None, 0, a, 1) == inval
    assert lib.dpc_server_read_matrices(a, 0, None, 1) == inval
    assert b"null" in lib.dpc_last_error()

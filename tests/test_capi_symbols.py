"""The C-ABI library loads (no GPU needed) and exports every function that
include/dpcull_hip.h declares; the Python binding declares the same set."""
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
    assert lib.dpc_abi_version() == 1

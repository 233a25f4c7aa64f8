import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # The CPU oracle is a few hundred lines of C; build it when it is missing
    # (it is test infrastructure, compiled with gcc -ffp-contract=off).
    if not (ROOT / "oracle" / "liboracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def lib():
    from pipeline_amd import _lib
    return _lib.load()


@pytest.fixture(autouse=True)
def _announce_gpu_case(request):
    # Print the case name before each GPU test runs, so a run that takes the
    # machine down still shows which case was running.
    if request.node.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu case] {request.node.nodeid}\n")
        sys.stdout.flush()
    yield


@pytest.fixture(params=[0, 16, 3], ids=["grid", "stream16", "stream3"])
def cull_shape(request):
    """Runs a test under both launch shapes of dpc_cull (full grid; persistent
    streaming grid, also with few waves so every wave walks many tiles)."""
    import ctypes as C
    from pipeline_amd import _lib
    lib = _lib.load()
    old = _lib.CullConfig()
    _lib.check(lib.dpc_cull_get_config(C.byref(old)))
    _lib.check(lib.dpc_cull_set_config(C.byref(_lib.CullConfig(request.param))))
    yield request.param
    _lib.check(lib.dpc_cull_set_config(C.byref(old)))

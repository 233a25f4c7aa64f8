"""Runs the C++ tests of the host mirrors (tests/cpp): the transform tree's
bookkeeping against the reference's Tree trace (CPU), the propagation plan's
tile runs against an emulation (CPU), the GPU tree, and the HIP culling
manager against the CPU manager restatement (GPU)."""
import subprocess

import pytest

from conftest import GOLDEN, ROOT

BIN = ROOT / "tests" / "cpp" / "bin"


def run(name, *args, timeout=300):
    exe = BIN / name
    if not exe.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True)
    res = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout, cwd=str(ROOT))
    print(res.stdout[-4000:])
    assert res.returncode == 0, res.stdout[-4000:] + res.stderr[-2000:]
    return res.stdout


def test_tree_layout_matches_reference_trace():
    out = run("tree_layout_test", str(GOLDEN / "ref_tree_trace.bin"))
    assert "0 failures" in out


def test_tree_plan_tiles_emulated():
    """The propagation plan's tile runs pass their consistency check, and an
    emulation of the tile kernel from their descriptors equals pass-by-pass
    propagation (random, index-reuse and adversarial trees, the 1M-node tree)."""
    out = run("tree_plan_test", timeout=600)
    assert "tree plan: 0 failures" in out


@pytest.mark.gpu
def test_gpu_tree_matches_reference_trace():
    out = run("tree_gpu_test", str(GOLDEN / "ref_tree_trace.bin"))
    assert "tree gpu: 0 failures" in out


@pytest.mark.gpu
def test_manager_matches_cpu_manager():
    out = run("manager_parity")
    assert "manager parity: 0 failures" in out


@pytest.mark.gpu
def test_xbar_integration_frames():
    """The reference-interface adapter (integration/dpcull_adapter.hpp) over the
    HIP manager, driven with xbar CullingImpl's call sequence and the tree
    observer that forwards changed world matrices to groupMatrixChanged, against
    the CPU manager restatement and a full cull of the current scene."""
    out = run("integration_test")
    assert "integration: 0 failures" in out


def test_host_code_under_sanitizers():
    """The tile planner (tree_plan.hpp), the tree bookkeeping (TreeLayout in
    host/Tree.cpp) and the serving loop's host code (host/Serving.cpp, with the
    concurrent callers of serve_stress.cpp against the host-only runtime of
    tests/cpp/host_runtime) built with AddressSanitizer and
    UndefinedBehaviorSanitizer (g++, host code only) and run: random,
    index-reuse and adversarial trees, the reference's recorded Tree trace, and
    serving batches whose deltas are gathered on one and on several threads.
    Any report fails the run."""
    import os
    subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp"), "sanitize"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    for args in (["asan_tree_plan_test", "--small"], ["asan_tree_layout_test", str(GOLDEN / "ref_tree_trace.bin")],
                 ["asan_serve_host"]):
        res = subprocess.run([str(BIN / args[0]), *args[1:]], capture_output=True, text=True, timeout=600,
                             cwd=str(ROOT), env=env)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
        assert "0 failures" in res.stdout
        assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr, res.stderr[-4000:]


@pytest.mark.gpu
def test_host_mirrors_recover_from_failing_hip_calls():
    """dpc_debug_inject_error makes the k-th checked HIP call fail: for every k
    inside a Tree::compute with a relayout and inside a Manager::cull, the call
    throws, and the retry gives the results of a run that never failed (world
    matrices vs the oracle; changed lists and visibility vs the CPU manager)."""
    out = run("fault_test")
    assert "fault injection: 0 failures" in out


@pytest.mark.gpu
def test_serve_concurrent_callers():
    """The serving loop with three submitter threads (each owning some scenes,
    waiting at once or later) and a statistics poller beside the scheduler and
    the per-stream issuers: every request succeeds and its masks equal those of
    the same requests run one at a time, for copied and directly read deltas."""
    out = run("serve_stress")
    assert "serve stress: 0 failures" in out


def test_serving_threads_under_thread_sanitizer():
    """The serving loop's host code (host/Serving.cpp: scheduler, per-stream
    issuers, callers' submit/wait, a statistics poller) built with
    ThreadSanitizer against a host-only runtime (tests/cpp/host_runtime: no
    device, asynchronous calls complete at once, events a little later) and run
    with the concurrent callers of serve_stress.cpp, on the CPU: every request
    succeeds and ThreadSanitizer reports no data race."""
    import os
    subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp"), "tsan"], check=True, timeout=600)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    res = subprocess.run([str(BIN / "tsan_serve_host")], capture_output=True, text=True, timeout=600,
                         cwd=str(ROOT), env=env)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-6000:]
    assert "serve stress: 0 failures" in res.stdout
    assert "ThreadSanitizer" not in res.stderr, res.stderr[-6000:]

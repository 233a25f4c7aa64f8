"""The allocation-free result diff (dpc_result_diff_noalloc) against the
workspace form (dpc_result_diff) through the C ABI: two calls in a row on the
same device, the second with a smaller input, so the second call runs over the
scratch contents the first one left.  Both must agree bit for bit, and with
the oracle (ResultBitSet::updateChanged, dp/culling/src/ResultBitSet.cpp:61-108)."""
import numpy as np
import pytest

from pipeline_amd import _lib
from pipeline_amd import device as dev

pytestmark = pytest.mark.gpu


def masks(n, seed, flip_rate):
    rng = np.random.default_rng(seed)
    words = (n + 63) // 64
    old = rng.integers(0, 2 ** 63, words, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, words, dtype=np.uint64)
    flip = rng.random(words * 64) < flip_rate
    new = old ^ np.packbits(flip, bitorder="little").view(np.uint64)
    if n % 64:
        keep = np.uint64((1 << (n % 64)) - 1)
        old[-1] &= keep
        new[-1] &= keep
    return old, new


def run_diff(n, old, new, noalloc):
    """One diff through the C ABI; returns (count, changed indices, old afterwards)."""
    lib = _lib.load()
    d_old = dev.DeviceArray.from_host(old)
    d_new = dev.DeviceArray.from_host(new)
    d_idx = dev.DeviceArray(max(n, 1), np.uint32)
    d_cnt = dev.DeviceArray(1, np.uint32)
    ws = None
    try:
        if noalloc:
            _lib.check(lib.dpc_result_diff_noalloc(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, 1, None))
        else:
            ws = dev.workspace(lib.dpc_result_diff_workspace(n))
            _lib.check(lib.dpc_result_diff(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, 1, ws.ptr, None))
        cnt = int(d_cnt.download()[0])
        return cnt, d_idx.download()[:cnt].copy(), d_old.download()
    finally:
        for d in (d_old, d_new, d_idx, d_cnt, ws):
            if d is not None:
                d.free()


# the first call takes the three-launch path (more than 1024 blocks of 256
# words), the second the two-launch path over a smaller mask
@pytest.mark.parametrize("sizes", [(20_000_000, 1_000_003), (2_000_000, 65)])
def test_noalloc_matches_workspace_form_on_two_calls(orc, sizes):
    for k, n in enumerate(sizes):
        old, new = masks(n, 100 + k, 0.01 if k == 0 else 0.3)
        want = orc.result_diff(n, old, new)
        got_ws = run_diff(n, old, new, noalloc=False)
        got_na = run_diff(n, old, new, noalloc=True)
        assert got_na[0] == got_ws[0] == want.size, (n, got_na[0], got_ws[0], want.size)
        assert np.array_equal(got_na[1], got_ws[1])
        assert np.array_equal(got_na[1], want)
        assert np.array_equal(got_na[2], got_ws[2]) and np.array_equal(got_na[2], new)


def test_noalloc_empty_and_size_limit():
    lib = _lib.load()
    d_cnt = dev.DeviceArray.from_host(np.array([7], np.uint32))
    try:
        _lib.check(lib.dpc_result_diff_noalloc(None, None, 0, None, d_cnt.ptr, 1, None))
        assert int(d_cnt.download()[0]) == 0
    finally:
        d_cnt.free()
    a = 0x10000  # never dereferenced: the size check comes first
    assert lib.dpc_result_diff_noalloc(a, a, (1 << 30) + 1, a, a, 1, None) == _lib.DPC_ERR_INVALID_ARGUMENT
    assert b"exceed" in lib.dpc_last_error()

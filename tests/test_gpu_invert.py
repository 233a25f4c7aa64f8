"""GPU parity of the world-matrix inverse transpose (dpc_inverse_transpose) with
dp::math::invertTranspose (dp/math/Matmnt.h:1171-1201): the golden inputs
written by the compiled reference (world-like chains, wide-range entries,
singular, non-finite and underflowing matrices) give the same success flags and
the same bits (any NaN equals any NaN); only the slots whose dirty byte is set
are written."""
import numpy as np
import pytest

from conftest import GOLDEN
from pipeline_amd import _lib
from pipeline_amd import device as dev

pytestmark = pytest.mark.gpu


def same_bits_or_both_nan(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


def run(m, dirty=None, fill=0.0):
    lib = _lib.load()
    n = m.shape[0]
    ld = n + 5  # planes wider than n: the kernel must use ld, not n
    src = dev.DeviceArray.from_host(dev.matrices_to_planes(m.reshape(-1, 4, 4), ld))
    out_host = np.full((4, ld, 4), fill, np.float32)
    out = dev.DeviceArray.from_host(out_host)
    ok = dev.DeviceArray(((n + 3) // 4) * 4, np.uint8)
    ok.fill_u32(0x07070707)
    d_dirty = dev.DeviceArray.from_host(dirty) if dirty is not None else None
    _lib.check(lib.dpc_inverse_transpose(src.ptr, ld, d_dirty.ptr if d_dirty else None, n, out.ptr, ld, ok.ptr, None))
    _lib.check(lib.dpc_stream_sync(None))
    res = np.transpose(out.download(), (1, 0, 2))[:n].reshape(n, 16)
    flags = ok.download()[:n]
    for a in (src, out, ok) + ((d_dirty,) if d_dirty else ()):
        a.free()
    return res, flags


def test_inverse_transpose_golden(orc):
    fx = np.load(GOLDEN / "ref_invert.npz")
    m = fx["m"]
    got, ok = run(m, fill=np.nan)
    assert np.array_equal(ok, fx["ok"])
    good = ok == 1
    assert same_bits_or_both_nan(got[good], fx["out"][good]).all()
    assert np.isnan(got[~good]).all()  # singular: output left as it was
    # and the oracle on the same inputs
    ok2, want = orc.invert_transpose(m)
    assert np.array_equal(ok2, ok)
    assert same_bits_or_both_nan(got[good], want.reshape(-1, 16)[good]).all()


def test_inverse_transpose_dirty_subset(orc):
    rng = np.random.default_rng(3)
    from pipeline_amd import scenes
    m = scenes.trs_matrices(rng, 5000, scale=(0.8, 1.25), trans=(-10, 10)).reshape(-1, 16)
    dirty = (rng.random(5000) < 0.3).astype(np.uint8)
    got, ok = run(m, dirty=dirty, fill=-1.0)
    _, want = orc.invert_transpose(m)
    d = dirty == 1
    assert (ok[d] == 1).all() and (ok[~d] == 0x07).all()  # clean slots: flag and output untouched
    assert np.array_equal(got[d].view(np.uint32), want.reshape(-1, 16)[d].view(np.uint32))
    assert (got[~d] == -1.0).all()


def test_inverse_transpose_rejects_bad_arguments(lib):
    a = dev.DeviceArray((4, 8, 4), np.float32)
    with pytest.raises(_lib.DpcError):
        _lib.check(lib.dpc_inverse_transpose(a.ptr, 8, None, 9, a.ptr, 8, None, None))  # n > ld
    with pytest.raises(_lib.DpcError):
        _lib.check(lib.dpc_inverse_transpose(None, 8, None, 4, a.ptr, 8, None, None))
    a.free()

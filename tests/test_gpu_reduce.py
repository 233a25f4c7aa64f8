"""GPU parity of the group bounding box (dpc_bbox, ManagerBitSet::calculateBoundingBox)
and of the result diff (dpc_result_diff, ResultBitSet::updateChanged) with the oracle."""
import ctypes as C

import numpy as np
import pytest

from pipeline_amd import _lib
from pipeline_amd import device as dev
from pipeline_amd import scenes

pytestmark = pytest.mark.gpu
F = np.float32


def gpu_bbox(lower, upper, mats, xform=None):
    lib = _lib.load()
    scene = dev.upload_scene(lower, upper, mats, xform)
    n = lower.shape[0]
    ws = dev.DeviceArray(max(lib.dpc_bbox_workspace(n), 4), np.uint8)
    out = dev.DeviceArray(6, F)
    _lib.check(lib.dpc_bbox(scene.box_a.ptr, scene.box_b.ptr, scene.xform_addr, n, scene.planes.ptr, scene.ld,
                            out.ptr, ws.ptr, None))
    r = out.download()
    for d in (ws, out):
        d.free()
    scene.free()
    return r


@pytest.mark.parametrize("n,seed", [(1, 1), (63, 2), (64, 3), (65, 4), (10_000, 5), (1_048_576, 6)])
def test_bbox_matches_oracle(orc, n, seed):
    rng = np.random.default_rng(seed)
    lower, upper = scenes.random_boxes(rng, n, (-1000.0, 1000.0))
    mats = scenes.trs_matrices(rng, n)
    got = gpu_bbox(lower, upper, mats)
    l4, e4 = scenes.extents(lower, upper)
    want = orc.bbox_group(l4, e4, np.arange(n), mats)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bbox_empty_and_signed_zero(orc):
    # an empty group returns (-FLT_MAX, +FLT_MAX) like the reference's Box3f(min, max)
    fmax = np.finfo(F).max
    lib = _lib.load()
    ws = dev.DeviceArray(max(lib.dpc_bbox_workspace(0), 4), np.uint8)
    out = dev.DeviceArray(6, F)
    dummy = dev.DeviceArray(64, F)
    _lib.check(lib.dpc_bbox(dummy.ptr, dummy.ptr, None, 0, dummy.ptr, 1, out.ptr, ws.ptr, None))
    assert np.array_equal(out.download(), np.array([-fmax] * 3 + [fmax] * 3, F))
    for d in (ws, out, dummy):
        d.free()
    # first-seen among +0/-0 survives, across objects far apart in the array
    ID = np.eye(4, dtype=F)
    mb = ID.copy()
    mb[3, 0] = -0.0
    n = 5000
    lower = np.tile(np.array([[5, 5, 5]], F), (n, 1))
    upper = lower + 1
    lower[17] = (0, 0, 0)
    upper[17] = (1, 1, 1)
    lower[4000] = (-0.0, -1, -1)
    upper[4000] = (1, 0, 0)
    xform = np.zeros(n, np.uint32)
    xform[4000] = 1
    mats = np.stack([ID, mb])
    got = gpu_bbox(lower, upper, mats, xform)
    l4, e4 = scenes.extents(lower, upper)
    want = orc.bbox_group(l4, e4, xform, mats)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got[0] == 0 and not np.signbit(got[0])
    lower[[17, 4000]] = lower[[4000, 17]]
    upper[[17, 4000]] = upper[[4000, 17]]
    xform[[17, 4000]] = xform[[4000, 17]]
    got = gpu_bbox(lower, upper, mats, xform)
    assert got[0] == 0 and np.signbit(got[0])


@pytest.mark.parametrize("n,ties,seed", [(300_000, 200, 1), (300_000, 3, 2), (2_000_000, 50, 3), (5_000, 40, 4)])
def test_bbox_signed_zero_ties_across_lanes_and_waves(orc, n, ties, seed):
    """Many objects whose lowest x is exactly +0 or -0, scattered over the whole
    array (different lanes, waves and partial boxes): the box keeps the zero of
    the first such object in index order, as the reference's sequential update."""
    rng = np.random.default_rng(seed)
    ID = np.eye(4, dtype=F)
    mb = ID.copy()
    mb[3, 0] = -0.0  # keeps the sign of lower.x: ((lx + -0) + -0) + -0 = lx for lx = +-0
    lower = np.tile(np.array([[5, 5, 5]], F), (n, 1)) + rng.uniform(0, 100, (n, 3)).astype(F)
    upper = lower + 1
    at = np.sort(rng.choice(n, ties, replace=False))
    signs = rng.integers(0, 2, ties)
    lower[at] = np.stack([np.where(signs == 1, F(-0.0), F(0.0)), np.full(ties, -1, F), np.full(ties, -1, F)], 1)
    upper[at] = (1, 0, 0)
    xform = np.zeros(n, np.uint32)
    xform[at] = 1
    mats = np.stack([ID, mb])
    got = gpu_bbox(lower, upper, mats, xform)
    l4, e4 = scenes.extents(lower, upper)
    want = orc.bbox_group(l4, e4, xform, mats)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert got[0] == 0 and bool(np.signbit(got[0])) == bool(signs[0])


@pytest.mark.parametrize("n", [1, 31, 64, 65, 1000, 16_385, 2_000_000, 20_000_000])
def test_result_diff_matches_oracle(orc, n):
    lib = _lib.load()
    rng = np.random.default_rng(n)
    words = (n + 63) // 64
    old = rng.integers(0, 2 ** 63, words, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, words, dtype=np.uint64)
    new = old.copy()
    flip = rng.random(words * 64) < 0.01
    new ^= np.packbits(flip, bitorder="little").view(np.uint64)
    if n % 64:
        keep = np.uint64((1 << (n % 64)) - 1)
        old[-1] &= keep
        new[-1] &= keep
    d_old = dev.DeviceArray.from_host(old)
    d_new = dev.DeviceArray.from_host(new)
    d_idx = dev.DeviceArray(n, np.uint32)
    d_cnt = dev.DeviceArray(1, np.uint32)
    ws = dev.DeviceArray(lib.dpc_result_diff_workspace(n), np.uint8)
    _lib.check(lib.dpc_result_diff(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, 1, ws.ptr, None))
    cnt = int(d_cnt.download()[0])
    want = orc.result_diff(n, old, new)
    assert cnt == want.size
    assert np.array_equal(d_idx.download()[:cnt], want)
    assert np.array_equal(d_old.download(), new)   # old = new afterwards
    for d in (d_old, d_new, d_idx, d_cnt, ws):
        d.free()

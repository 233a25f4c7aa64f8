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
    ws = dev.workspace(lib.dpc_bbox_workspace(n))
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
    ws = dev.workspace(lib.dpc_bbox_workspace(0))
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
    ws = dev.workspace(lib.dpc_result_diff_workspace(n))
    _lib.check(lib.dpc_result_diff(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, 1, ws.ptr, None))
    cnt = int(d_cnt.download()[0])
    want = orc.result_diff(n, old, new)
    assert cnt == want.size
    assert np.array_equal(d_idx.download()[:cnt], want)
    assert np.array_equal(d_old.download(), new)   # old = new afterwards
    for d in (d_old, d_new, d_idx, d_cnt, ws):
        d.free()


def _diff_inputs(n, seed, rate=0.01):
    rng = np.random.default_rng(seed)
    words = (n + 63) // 64
    old = rng.integers(0, 2 ** 63, words, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, words, dtype=np.uint64)
    new = old ^ np.packbits(rng.random(words * 64) < rate, bitorder="little").view(np.uint64)
    if n % 64:
        keep = np.uint64((1 << (n % 64)) - 1)
        old[-1] &= keep
        new[-1] &= keep
    return old, new


def test_result_diff_one_workspace_many_calls(orc):
    """One workspace for a sequence of calls of different sizes (different
    workgroup counts): the look-back state each call leaves serves the next
    one, with no clearing in between."""
    lib = _lib.load()
    sizes = [20_000_000, 1_000, 2_000_000, 65, 20_000_000, 16_385, 4_000_000, 1]
    nmax = max(sizes)
    ws = dev.workspace(lib.dpc_result_diff_workspace(nmax))
    d_idx = dev.DeviceArray(nmax, np.uint32)
    d_cnt = dev.DeviceArray(1, np.uint32)
    for k, n in enumerate(sizes):
        old, new = _diff_inputs(n, 100 + k, 0.002 if n > 10_000_000 else 0.05)
        d_old = dev.DeviceArray.from_host(old)
        d_new = dev.DeviceArray.from_host(new)
        _lib.check(lib.dpc_result_diff(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, k % 2, ws.ptr, None))
        cnt = int(d_cnt.download()[0])
        want = orc.result_diff(n, old, new)
        assert cnt == want.size, (k, n, cnt, want.size)
        assert np.array_equal(d_idx.download()[:cnt], want), (k, n)
        assert np.array_equal(d_old.download(), new if k % 2 else old), (k, n)
        d_old.free()
        d_new.free()
    for d in (ws, d_idx, d_cnt):
        d.free()


def test_result_diff_graph_replays(orc):
    """The diff captured once in a hipGraph and replayed with new masks written
    into its inputs between replays: every replay equals the oracle (the look-back
    tag comes from device memory, not from the captured arguments)."""
    lib = _lib.load()
    n = 3_000_000
    stream = dev.Stream()
    old0, new0 = _diff_inputs(n, 7)
    d_old = dev.DeviceArray.from_host(old0)
    d_new = dev.DeviceArray.from_host(new0)
    d_idx = dev.DeviceArray(n, np.uint32)
    d_cnt = dev.DeviceArray(1, np.uint32)
    ws = dev.workspace(lib.dpc_result_diff_workspace(n), stream.handle)
    stream.sync()
    graph = dev.Graph(stream, lambda: _lib.check(lib.dpc_result_diff(d_old.ptr, d_new.ptr, n, d_idx.ptr, d_cnt.ptr, 0,
                                                                     ws.ptr, stream.handle)))
    for rep in range(5):
        old, new = _diff_inputs(n, 200 + rep, 0.001 * (rep + 1))
        d_old.upload(old, stream.handle)
        d_new.upload(new, stream.handle)
        graph.launch(stream)
        cnt = int(d_cnt.download(stream.handle)[0])
        want = orc.result_diff(n, old, new)
        assert cnt == want.size, (rep, cnt, want.size)
        assert np.array_equal(d_idx.download(stream.handle)[:cnt], want), rep
    graph.destroy()
    for d in (d_old, d_new, d_idx, d_cnt, ws):
        d.free()
    stream.destroy()


def test_bbox_one_workspace_many_calls_and_graph(orc):
    """One bounding-box workspace for calls over different object counts (so
    different workgroup counts), then a captured call replayed after the
    matrices move: every box equals the oracle's."""
    lib = _lib.load()
    rng = np.random.default_rng(12)
    sizes = [2_000_000, 70, 300_000, 1, 2_000_000, 5_000]
    ws = dev.workspace(max(lib.dpc_bbox_workspace(n) for n in sizes))
    out = dev.DeviceArray(6, F)
    for n in sizes:
        lower, upper = scenes.random_boxes(rng, n, (-1000.0, 1000.0))
        mats = scenes.trs_matrices(rng, n)
        scene = dev.upload_scene(lower, upper, mats)
        _lib.check(lib.dpc_bbox(scene.box_a.ptr, scene.box_b.ptr, scene.xform_addr, n, scene.planes.ptr, scene.ld,
                                out.ptr, ws.ptr, None))
        l4, e4 = scenes.extents(lower, upper)
        want = orc.bbox_group(l4, e4, np.arange(n), mats)
        assert np.array_equal(out.download().view(np.uint32), want.view(np.uint32)), n
        scene.free()
    n = 500_000
    lower, upper = scenes.random_boxes(rng, n, (-1000.0, 1000.0))
    mats = scenes.trs_matrices(rng, n)
    scene = dev.upload_scene(lower, upper, mats)
    stream = dev.Stream()
    graph = dev.Graph(stream, lambda: _lib.check(lib.dpc_bbox(scene.box_a.ptr, scene.box_b.ptr, scene.xform_addr, n,
                                                              scene.planes.ptr, scene.ld, out.ptr, ws.ptr,
                                                              stream.handle)))
    l4, e4 = scenes.extents(lower, upper)
    for rep in range(4):
        mats = scenes.trs_matrices(rng, n) * np.float32(1 + rep)
        planes = np.ascontiguousarray(np.transpose(mats, (1, 0, 2))).reshape(-1)  # row planes
        scene.planes.upload(planes, stream.handle)
        graph.launch(stream)
        got = out.download(stream.handle)
        want = orc.bbox_group(l4, e4, np.arange(n), mats)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), rep
    graph.destroy()
    stream.destroy()
    scene.free()
    for d in (ws, out):
        d.free()

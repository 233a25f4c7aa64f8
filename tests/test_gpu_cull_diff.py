"""GPU parity of the cull with the result diff fused into it (dpc_cull_diff: the
tail of cpu::ManagerImpl::cull, dp/culling/cpu/src/ManagerImpl.cpp:467-518,
which hands its mask to ResultBitSet::updateChanged,
dp/culling/src/ResultBitSet.cpp:61-108).  The mask must equal the oracle's
cull, and the changed indices, their count and the updated old mask the
oracle's result diff of that mask against the old one, bit for bit, under both
launch shapes of the cull, on both sides of the size up to which the count is
fused (2^22 objects), with garbage in the old mask's bits past n."""
import numpy as np
import pytest

from gpu_util import mismatches
from pipeline_amd import _lib
from pipeline_amd import device as dev
from pipeline_amd import scenes

pytestmark = pytest.mark.gpu
FUSED_MAX = 1 << 22


def _old_mask(rng, want, n, rate):
    """The previous frame's mask: the new one with a fraction of its bits flipped,
    plus set bits past n (which the diff must ignore)."""
    old = want ^ np.packbits(rng.random(want.size * 64) < rate, bitorder="little").view(np.uint64)
    if n % 64:
        old[-1] |= ~np.uint64((1 << (n % 64)) - 1)
    return old


def _oracle_diff(orc, n, old, new):
    clean = old.copy()
    if n % 64:
        clean[-1] &= np.uint64((1 << (n % 64)) - 1)
    return orc.result_diff(n, clean, new)


class _Call:
    """Device buffers of one dpc_cull_diff call over a scene of n objects."""

    def __init__(self, lib, n, ws=None):
        words = max(dev.mask_words(n), 1)
        self.mask = dev.DeviceArray(words, np.uint64)
        self.idx = dev.DeviceArray(max(n, 1), np.uint32)
        self.cnt = dev.DeviceArray(1, np.uint32)
        self.own_ws = ws is None
        self.ws = dev.workspace(lib.dpc_cull_diff_workspace(n)) if ws is None else ws

    def free(self):
        for d in (self.mask, self.idx, self.cnt) + ((self.ws,) if self.own_ws else ()):
            d.free()


def _run(lib, scene, vp, call, d_old, update_old, stream=None):
    _lib.check(lib.dpc_cull_diff(scene.box_a.ptr, scene.box_b.ptr, scene.xform_addr, scene.n_objects,
                                 scene.planes.ptr, scene.ld, dev.vp_arg(vp), call.mask.ptr, d_old.ptr, call.idx.ptr,
                                 call.cnt.ptr, update_old, call.ws.ptr, stream))


def _check(orc, n, call, d_old, old, want, update_old, label):
    got = call.mask.download()
    assert mismatches(got, want, n) == 0, label
    cnt = int(call.cnt.download()[0])
    expect = _oracle_diff(orc, n, old, want)
    assert cnt == expect.size, (label, cnt, expect.size)
    assert np.array_equal(call.idx.download()[:cnt], expect), label
    assert np.array_equal(d_old.download(), want if update_old else old), label


def _flat(orc, n, seed):
    s = scenes.flat_scene(n, seed, (-40.0, 40.0), 70.0, 1.3, 0.1, 80.0)
    l4, e4 = scenes.extents(s.lower, s.upper)
    want = orc.cull_group(l4, e4, np.arange(n), s.matrices, s.vp)
    return s, want


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4097, 100_000, 1_048_576, FUSED_MAX, FUSED_MAX + 1])
def test_cull_diff_matches_oracle(orc, lib, n, cull_shape):
    rng = np.random.default_rng(n)
    s, want = _flat(orc, n, 500 + n % 1000)
    old = _old_mask(rng, want, n, 0.01)
    scene = dev.upload_scene(s.lower, s.upper, s.matrices)
    d_old = dev.DeviceArray.from_host(old)
    call = _Call(lib, n)
    try:
        _run(lib, scene, s.vp, call, d_old, 1)
        _check(orc, n, call, d_old, old, want, 1, f"n={n}")
    finally:
        call.free()
        d_old.free()
        scene.free()


def test_cull_diff_shared_matrices_through_indices(orc, lib, cull_shape):
    """Objects reference matrices by transform index (four objects per matrix, as
    the leaves of a transform tree); the old mask is left as it was."""
    rng = np.random.default_rng(9)
    n, nm = 300_000, 75_000
    lower, upper = scenes.random_boxes(rng, n, (-20.0, 20.0))
    mats = scenes.trs_matrices(rng, nm)
    xform = (np.arange(n) // 4).astype(np.uint32)
    vp = scenes.perspective_vp([0, 0, 30], [0, 0, -1], 60, 1.5, 0.1, 100)
    l4, e4 = scenes.extents(lower, upper)
    want = orc.cull_group(l4, e4, xform, mats, vp)
    old = _old_mask(rng, want, n, 0.05)
    scene = dev.upload_scene(lower, upper, mats, xform)
    d_old = dev.DeviceArray.from_host(old)
    call = _Call(lib, n)
    try:
        _run(lib, scene, vp, call, d_old, 0)
        _check(orc, n, call, d_old, old, want, 0, "indexed")
    finally:
        call.free()
        d_old.free()
        scene.free()


def test_cull_diff_one_workspace_many_calls(orc, lib):
    """One workspace for calls of different sizes, fused and not: the per-word
    counts a call leaves behind never reach the next one."""
    sizes = [100_000, FUSED_MAX + 4_000, 1_000, FUSED_MAX, 65, 2_000_000]
    ws = dev.workspace(lib.dpc_cull_diff_workspace(max(sizes)))
    try:
        for k, n in enumerate(sizes):
            rng = np.random.default_rng(40 + k)
            s, want = _flat(orc, n, 40 + k)
            old = _old_mask(rng, want, n, 0.002 if n > 1_000_000 else 0.05)
            scene = dev.upload_scene(s.lower, s.upper, s.matrices)
            d_old = dev.DeviceArray.from_host(old)
            call = _Call(lib, n, ws)
            try:
                _run(lib, scene, s.vp, call, d_old, k % 2)
                _check(orc, n, call, d_old, old, want, k % 2, f"call {k}, n={n}")
            finally:
                call.free()
                d_old.free()
                scene.free()
    finally:
        ws.free()


def test_cull_diff_graph_replays(orc, lib):
    """The fused cull and diff captured once in a hipGraph and replayed with a new
    old mask uploaded between replays (eager copies and synchronisations in
    between): every replay equals the oracle."""
    n = 1_000_000
    s, want = _flat(orc, n, 77)
    stream = dev.Stream()
    scene = dev.upload_scene(s.lower, s.upper, s.matrices)
    d_old = dev.DeviceArray(max(dev.mask_words(n), 1), np.uint64)
    call = _Call(lib, n)
    stream.sync()
    graph = dev.Graph(stream, lambda: _run(lib, scene, s.vp, call, d_old, 0, stream.handle))
    try:
        for rep in range(4):
            old = _old_mask(np.random.default_rng(700 + rep), want, n, 0.001 * (rep + 1))
            d_old.upload(old, stream.handle)
            stream.sync()
            graph.launch(stream)
            stream.sync()
            _check(orc, n, call, d_old, old, want, 0, f"replay {rep}")
    finally:
        graph.destroy()
        call.free()
        d_old.free()
        scene.free()
        stream.destroy()

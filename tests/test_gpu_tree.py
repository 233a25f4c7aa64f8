"""GPU parity of level-synchronous world-matrix propagation (dpc_tree_propagate)
with the oracle of dp::transform::Tree::compute.  World matrices bit-exact
(0 ulp, well inside the 2-ulp bound), plus the FP64 error report; then the cull
over the propagated matrices, bit-exact."""
import numpy as np
import pytest

from gpu_util import mismatches
from pipeline_amd import _lib
from pipeline_amd import device as dev
from pipeline_amd import scenes

pytestmark = pytest.mark.gpu
F = np.float32


def ulp_diff(a, b):
    ia = np.ascontiguousarray(a, F).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, F).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


class DeviceTree:
    def __init__(self, ts: scenes.TreeScene):
        self.ts = ts
        self.n = ts.n_slots
        self.local = dev.DeviceArray.from_host(dev.matrices_to_planes(ts.local))
        world0 = np.zeros_like(ts.local)
        world0[0] = np.eye(4, dtype=F)
        self.world = dev.DeviceArray.from_host(dev.matrices_to_planes(world0))
        self.parent = dev.DeviceArray.from_host(ts.parent_slot)
        self.dirty_world = dev.DeviceArray(((self.n + 3) // 4) * 4, np.uint8)
        self.dirty_world.fill_u32(0)
        self.offsets = np.ascontiguousarray(ts.level_offsets, np.uint64)

    def plan(self, tile_slots=0):
        """dpc_tree_plan_create on this layout: (handle, launches, tile runs, tiles)."""
        import ctypes as C
        lib = _lib.load()
        h = C.c_void_p()
        par = np.ascontiguousarray(self.ts.parent_slot, np.uint32)
        _lib.check(lib.dpc_tree_plan_create(par.ctypes.data, self.n, self.offsets.ctypes.data, self.offsets.size - 1,
                                            tile_slots, C.byref(h)))
        launches, runs, tiles = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _lib.check(lib.dpc_tree_plan_info(h, C.byref(launches), C.byref(runs), C.byref(tiles)))
        return h, launches.value, runs.value, tiles.value

    def propagate(self, dirty_local=None, stream=None, plan=None):
        import ctypes as C
        lib = _lib.load()
        d_dl = None
        if dirty_local is not None:
            d_dl = dev.DeviceArray.from_host(dirty_local)
        if plan is not None:
            _lib.check(lib.dpc_tree_propagate_plan(plan, self.local.ptr, self.world.ptr, d_dl.ptr if d_dl else None,
                                                   self.dirty_world.ptr, stream))
        else:
            off = (C.c_uint64 * self.offsets.size)(*self.offsets.tolist())
            _lib.check(lib.dpc_tree_propagate(self.local.ptr, self.world.ptr, self.n, self.parent.ptr, off,
                                              self.offsets.size - 1, d_dl.ptr if d_dl else None,
                                              self.dirty_world.ptr, stream))
        _lib.check(lib.dpc_stream_sync(stream))
        if d_dl:
            d_dl.free()

    def world_by_slot(self):
        return np.transpose(self.world.download(), (1, 0, 2))

    def free(self):
        for a in (self.local, self.world, self.parent, self.dirty_world):
            a.free()


def oracle_world(orc, ts):
    """Tree::compute on the reference's index-addressed arrays and level lists."""
    offsets, entries = ts.reference_levels()
    local = np.ascontiguousarray(ts.local_by_index().reshape(-1, 16))
    world = np.zeros_like(local)
    world[0] = np.eye(4, dtype=F).reshape(16)
    dirty_local = np.ones(ts.n_slots, np.uint8)
    dirty_world = np.zeros(ts.n_slots, np.uint8)
    event = orc.tree_compute(offsets, entries, local, world, dirty_local, dirty_world)
    return world.reshape(-1, 4, 4), event


@pytest.mark.parametrize("n_nodes,seed,kw", [
    (37, 1, {}), (2_000, 5, {}), (50_000, 8, {}),
    # two levels above the single-workgroup size, then a small one whose parents
    # were written by an earlier launch
    (6_500, 9, {"roots": 3_000, "p": 0.99}),
    # deeper than one tile run (several runs, each first pass reading parents
    # written by the previous launch), a flat two-level tree of 100k + 100k
    # nodes, and a tree that grows from a single top-level transform
    (4_000, 11, {"max_depth": 120, "p": 0.97, "max_children": 2, "roots": 30}),
    (6_000, 14, {"max_depth": 220, "p": 0.995, "max_children": 2, "roots": 16}),
    (200_000, 12, {"max_depth": 2, "roots": 100_000, "p": 0.9}),
    (60_000, 13, {"max_depth": 40, "p": 0.5, "max_children": 8, "roots": 1}),
])
@pytest.mark.parametrize("use_plan", [None, 1024, 512, 256], ids=["levels", "plan1024", "plan512", "plan256"])
def test_propagate_matches_oracle(orc, n_nodes, seed, kw, use_plan):
    ts = scenes.tree_scene(n_nodes, seed, **kw)
    t = DeviceTree(ts)
    plan = t.plan(use_plan)[0] if use_plan else None
    t.propagate(plan=plan)
    if plan is not None:
        _lib.load().dpc_tree_plan_destroy(plan)
    got = t.world_by_slot()
    want, event = oracle_world(orc, ts)
    want_by_slot = want[ts.index_of_slot]
    assert np.array_equal(got[1:].view(np.uint32), want_by_slot[1:].view(np.uint32))
    dw = t.dirty_world.download()[: ts.n_slots]
    assert np.array_equal(dw[1:], event[ts.index_of_slot][1:])
    t.free()


def test_tree_1m_and_cull(orc):
    """1,000,000 nodes (depth <= 16, geometric children), then cull of ~2.4M leaf objects."""
    ts = scenes.tree_scene(1_000_000, 2)
    t = DeviceTree(ts)
    t.propagate()
    got = t.world_by_slot()
    want, _ = oracle_world(orc, ts)
    want_by_slot = want[ts.index_of_slot]
    ulps = ulp_diff(got[1:], want_by_slot[1:])
    # FP64 oracle of the same recurrence: max relative error of the FP32 result
    offsets, entries = ts.reference_levels()
    w64 = orc.tree_compute_f64(offsets, entries, ts.local_by_index().astype(np.float64))[ts.index_of_slot]
    scale = np.abs(w64[1:]).max(axis=(1, 2), keepdims=True)
    rel = float((np.abs(got[1:].astype(np.float64) - w64[1:]) / scale).max())
    print(f"tree 1M: {ts.n_slots - 1} nodes, {len(ts.level_offsets) - 1} levels, max ulp vs FP32 oracle "
          f"{int(ulps.max())}, max relative error vs FP64 {rel:.3e}")
    assert ulps.max() == 0
    assert rel < 1e-4

    # cull the leaf objects against the propagated world matrices (by slot)
    box = dev.upload_scene(ts.obj_lower, ts.obj_upper, np.eye(4, dtype=F)[None], None)
    box.planes.free()
    box.planes, box.ld = t.world, ts.n_slots
    box.xform = dev.DeviceArray.from_host(ts.obj_slot)
    mask = dev.cull(box, ts.vp).download()
    l4, e4 = scenes.extents(ts.obj_lower, ts.obj_upper)
    want_mask = orc.cull_group(l4, e4, ts.obj_slot, want_by_slot, ts.vp)
    m = ts.obj_slot.shape[0]
    n_mis = mismatches(mask, want_mask, m)
    print(f"tree 1M cull: {m} objects, {int(np.unpackbits(want_mask.view(np.uint8)).sum())} visible, "
          f"{n_mis} mismatching bits")
    assert n_mis == 0
    box.box_a.free()
    box.box_b.free()
    box.xform.free()
    t.free()


def test_dirty_propagation(orc):
    """Second compute with a few dirty locals: only dirty nodes and their
    descendants are recomputed (Tree.cpp:155-159), and the event lists exactly them."""
    ts = scenes.tree_scene(20_000, 17)
    t = DeviceTree(ts)
    t.propagate()
    t.dirty_world.fill_u32(0)   # the caller clears after consuming the event (Tree.cpp:164-165)
    # oracle state after the first compute
    offsets, entries = ts.reference_levels()
    local = np.ascontiguousarray(ts.local_by_index().reshape(-1, 16))
    world = np.zeros_like(local)
    world[0] = np.eye(4, dtype=F).reshape(16)
    orc.tree_compute(offsets, entries, local, world, np.ones(ts.n_slots, np.uint8), np.zeros(ts.n_slots, np.uint8))
    # change some locals (by slot on the device, by index in the oracle)
    rng = np.random.default_rng(3)
    slots = rng.choice(np.arange(1, ts.n_slots), 40, replace=False)
    new = scenes.trs_matrices(rng, 40, scale=(0.8, 1.25), trans=(-10, 10))
    import ctypes as C
    lib = _lib.load()
    d_idx = dev.DeviceArray.from_host(slots.astype(np.uint32))
    d_new = dev.DeviceArray.from_host(new.reshape(-1, 16))
    _lib.check(lib.dpc_matrices_scatter(t.local.ptr, t.n, d_idx.ptr, d_new.ptr, 40, None))
    dirty_slots = np.zeros(ts.n_slots, np.uint8)
    dirty_slots[slots] = 1
    t.propagate(dirty_local=dirty_slots)
    idx = ts.index_of_slot[slots]
    local[idx] = new.reshape(-1, 16)
    dl = np.zeros(ts.n_slots, np.uint8)
    dl[idx] = 1
    event = orc.tree_compute(offsets, entries, local, world, dl, np.zeros(ts.n_slots, np.uint8))
    got = t.world_by_slot()
    want = world.reshape(-1, 4, 4)[ts.index_of_slot]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    dw = t.dirty_world.download()[: ts.n_slots]
    assert np.array_equal(dw, event[ts.index_of_slot])
    assert 40 <= dw.sum() < ts.n_slots // 2
    # the event as a BitArray over transform indices
    bits = dev.DeviceArray(((ts.n_slots + 63) // 64) * 2, np.uint32)
    bits.fill_u32(0)
    d_map = dev.DeviceArray.from_host(ts.index_of_slot)
    _lib.check(lib.dpc_bytes_to_bits_indexed(t.dirty_world.ptr, d_map.ptr, ts.n_slots, bits.ptr, ts.n_slots, None))
    got_bits = np.unpackbits(bits.download().view(np.uint8), bitorder="little")[: ts.n_slots]
    assert np.array_equal(got_bits, event)
    for d in (d_idx, d_new, bits, d_map):
        d.free()
    t.free()


@pytest.mark.parametrize("tile_slots", [1024, 512, 256])
def test_plan_tiles_tree_1m_and_dirty_frames(orc, tile_slots):
    """The 1,000,000-node tree through the plan: its deep levels form a tile run;
    a full compute and then two partial ones (a few hundred dirty locals) give
    the oracle's world matrices and events bit for bit."""
    ts = scenes.tree_scene(1_000_000, 2)
    t = DeviceTree(ts)
    plan, launches, runs, tiles = t.plan(tile_slots)
    print(f"plan: {launches} launches, {runs} tile runs, {tiles} tiles for {len(ts.level_offsets) - 1} passes")
    assert runs >= 1 and launches < len(ts.level_offsets) - 1
    offsets, entries = ts.reference_levels()
    local = np.ascontiguousarray(ts.local_by_index().reshape(-1, 16))
    world = np.zeros_like(local)
    world[0] = np.eye(4, dtype=F).reshape(16)
    rng = np.random.default_rng(11)
    import ctypes as C
    lib = _lib.load()
    for frame in range(3):
        if frame == 0:
            dl_slots = np.ones(ts.n_slots, np.uint8)
            t.propagate(plan=plan)
        else:
            slots = rng.choice(np.arange(1, ts.n_slots), 300, replace=False)
            new = scenes.trs_matrices(rng, 300, scale=(0.8, 1.25), trans=(-10, 10))
            d_idx = dev.DeviceArray.from_host(slots.astype(np.uint32))
            d_new = dev.DeviceArray.from_host(new.reshape(-1, 16))
            _lib.check(lib.dpc_matrices_scatter(t.local.ptr, t.n, d_idx.ptr, d_new.ptr, 300, None))
            d_idx.free()
            d_new.free()
            dl_slots = np.zeros(ts.n_slots, np.uint8)
            dl_slots[slots] = 1
            local[ts.index_of_slot[slots]] = new.reshape(-1, 16)
            t.propagate(dirty_local=dl_slots, plan=plan)
        dl = np.zeros(ts.n_slots, np.uint8)
        dl[ts.index_of_slot] = dl_slots
        event = orc.tree_compute(offsets, entries, local, world, dl, np.zeros(ts.n_slots, np.uint8))
        got = t.world_by_slot()
        want = world.reshape(-1, 4, 4)[ts.index_of_slot]
        assert np.array_equal(got[1:].view(np.uint32), want[1:].view(np.uint32)), f"frame {frame}"
        dw = t.dirty_world.download()[: ts.n_slots]
        assert np.array_equal(dw[1:], event[ts.index_of_slot][1:]), f"frame {frame}"
        t.dirty_world.fill_u32(0)
    lib.dpc_tree_plan_destroy(plan)
    t.free()


def test_plan_rejects_bad_arguments():
    import ctypes as C
    lib = _lib.load()
    h = C.c_void_p()
    par = np.zeros(4, np.uint32)
    off = np.array([1, 3, 5], np.uint64)  # ends beyond ld = 4
    with pytest.raises(_lib.DpcError):
        _lib.check(lib.dpc_tree_plan_create(par.ctypes.data, 4, off.ctypes.data, 2, 0, C.byref(h)))
    off = np.array([1, 3, 4], np.uint64)
    with pytest.raises(_lib.DpcError):  # an unsupported tile width
        _lib.check(lib.dpc_tree_plan_create(par.ctypes.data, 4, off.ctypes.data, 2, 300, C.byref(h)))
    _lib.check(lib.dpc_tree_plan_create(par.ctypes.data, 4, off.ctypes.data, 2, 0, C.byref(h)))
    lib.dpc_tree_plan_destroy(h)

"""GPU parity of batched multi-frustum culling (dpc_cull_batch, one persistent
launch) and of its per-view form (dpc_cull_per_view, one launch per frustum):
every (frustum, group) mask equals the oracle's cull(group, frustum), bit for bit,
for every tile/wave shape, with and without hipGraph capture and replay."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import mismatches
from pipeline_amd import _lib
from pipeline_amd import device as dev
from pipeline_amd import scenes

pytestmark = pytest.mark.gpu
F = np.float32


def oracle_pairs(orc, bs: scenes.BatchedScene) -> np.ndarray:
    """Expected mask words for every pair, laid out at the pairs' offsets."""
    out = np.zeros(max(bs.mask_words, 1), np.uint64)
    l4, e4 = scenes.extents(bs.lower, bs.upper)
    for f, g, off in zip(bs.pair_frustum, bs.pair_group, bs.pair_offset):
        a, n = int(bs.group_first[g]), int(bs.group_count[g])
        m = orc.cull_group(l4[a:a + n], e4[a:a + n], np.arange(a, a + n), bs.matrices, bs.vps[f])
        out[int(off):int(off) + m.size] = m
    return out


class Batch:
    def __init__(self, bs, cfg=None):
        self.bs = bs
        self.scene = dev.upload_scene(bs.lower, bs.upper, bs.matrices)
        self.plan = dev.batched_plan(bs, cfg)
        self.plan.bind([self.scene])
        assert self.plan.mask_words == bs.mask_words and self.plan.tests == bs.tests
        self.vps = dev.DeviceArray.from_host(bs.vps.reshape(-1, 16))
        self.masks = dev.DeviceArray(max(bs.mask_words, 1), np.uint64)

    def run_batch(self, stream=None):
        self.plan.cull(self.vps, self.masks, stream)

    def run_per_view(self, stream=None):
        self.plan.cull_per_view(self.bs.vps, self.masks, stream)

    def result(self):
        return self.masks.download()

    def free(self):
        self.plan.destroy()
        self.scene.free()
        self.vps.free()
        self.masks.free()


def small_scene(seed=7):
    return scenes.batched_scene(seed=seed, n=120_000, n_groups=12, n_cascades=2, n_cube=3, n_cameras=7,
                                size_range=(100, 40_000), groups_per_frustum=(1, 6), extent=200.0)


CONFIGS = [(1, 256, 8, 0), (2, 256, 8, 0), (4, 256, 4, 0), (2, 128, 16, 4), (1, 64, 32, 1), (4, 64, 8, 3)]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_batch_matches_oracle_all_shapes(orc, cfg):
    bs = small_scene()
    want = oracle_pairs(orc, bs)
    b = Batch(bs, _lib.BatchConfig(*cfg))
    b.masks.fill_u32(0xDEADBEEF)
    b.run_batch()
    got = b.result()
    assert mismatches(got, want, bs.mask_words * 64) == 0
    b.free()


def test_per_view_matches_oracle(orc):
    bs = small_scene(8)
    want = oracle_pairs(orc, bs)
    b = Batch(bs)
    b.masks.fill_u32(0xDEADBEEF)
    b.run_per_view()
    assert mismatches(b.result(), want, bs.mask_words * 64) == 0
    b.free()


def test_batch_equals_per_view_2m_objects_64_frusta(orc):
    """Full-size batched workload: 2,097,152 objects, 64 groups, 64 frusta."""
    bs = scenes.batched_scene()
    b = Batch(bs)
    b.run_batch()
    batched = b.result()
    b.masks.fill_u32(0)
    b.run_per_view()
    per_view = b.result()
    assert np.array_equal(batched, per_view)
    want = oracle_pairs(orc, bs)
    n_mis = mismatches(batched, want, bs.mask_words * 64)
    print(f"batched 2M x 64: {bs.n} objects, {len(bs.pair_group)} pairs, {bs.tests} tests, "
          f"{int(np.unpackbits(want.view(np.uint8)).sum())} visible, {n_mis} mismatching bits")
    assert n_mis == 0
    b.free()


def test_degenerate_matrices_and_frusta(orc):
    """Inputs off the common path: waves with one projective object matrix, a
    whole group of them, last columns of -0, M33 = 2, zero extents, and frusta
    whose row 3 holds inf, NaN, -inf or a huge finite value (some in the middle
    of an item's frusta).  Batched launches of several shapes and the per-view
    launch all equal the oracle bit for bit."""
    bs = small_scene(13)
    rng = np.random.default_rng(5)
    n = bs.n
    m = bs.matrices
    proj = rng.choice(n, n // 500, replace=False)
    m[proj, 0:3, 3] = rng.uniform(-0.01, 0.01, (proj.size, 3)).astype(F)
    negz = rng.choice(n, n // 50, replace=False)
    m[negz, 0:3, 3] = F(-0.0)
    w2 = rng.choice(n, n // 1000, replace=False)
    m[w2, 3, 3] = F(2.0)
    flat = rng.choice(n, n // 100, replace=False)
    bs.upper[flat] = bs.lower[flat]
    # one group entirely projective: all of its waves take the full test
    g = int(np.argmax(bs.group_count))
    a, c = int(bs.group_first[g]), int(bs.group_count[g])
    m[a:a + c, 2, 3] = F(1e-3)
    nf = bs.vps.shape[0]
    bs.vps[1 % nf, 3, 1] = np.inf
    bs.vps[3 % nf, 3, 2] = np.nan
    bs.vps[5 % nf, 3, 0] = F(3e38)
    bs.vps[7 % nf, 3, 3] = -np.inf
    want = oracle_pairs(orc, bs)
    for cfg in [None, (1, 256, 8, 0), (2, 128, 16, 4), (4, 64, 8, 3)]:
        b = Batch(bs, None if cfg is None else _lib.BatchConfig(*cfg))
        b.masks.fill_u32(0xDEADBEEF)
        b.run_batch()
        assert mismatches(b.result(), want, bs.mask_words * 64) == 0, f"batched {cfg}"
        if cfg is None:
            b.masks.fill_u32(0xDEADBEEF)
            b.run_per_view()
            assert mismatches(b.result(), want, bs.mask_words * 64) == 0, "per view"
        b.free()


def test_graph_replay_with_moving_frusta(orc):
    """The batched launch captured once in a hipGraph and replayed while the
    frusta change in device memory between replays (a camera moving per frame):
    every replay equals the eager result and the oracle."""
    bs = small_scene(9)
    b = Batch(bs)
    stream = dev.Stream()
    graph = dev.Graph(stream, lambda: b.run_batch(stream.handle))
    rng = np.random.default_rng(1)
    for rep in range(4):
        if rep:
            jitter = scenes.perspective_vp(rng.uniform(-50, 50, 3), rng.normal(size=3), 60, 1.5, 0.1, 300)
            bs.vps[rep % bs.vps.shape[0]] = jitter
            b.vps.upload(bs.vps.reshape(-1, 16))
        b.masks.fill_u32(0)
        graph.launch(stream)
        stream.sync()
        replay = b.result()
        b.masks.fill_u32(0)
        b.run_batch()
        _lib.check(_lib.load().dpc_stream_sync(None))
        eager = b.result()
        assert np.array_equal(replay, eager), f"replay {rep}"
        assert mismatches(replay, oracle_pairs(orc, bs), bs.mask_words * 64) == 0, f"replay {rep}"
    graph.destroy()
    stream.destroy()
    b.free()


def test_bind_and_launch_reject_bad_extents():
    """Group ranges beyond a bound scene, an unbound plan and short mask buffers
    are errors reported before any launch."""
    import ctypes as C
    lib = _lib.load()
    bs = small_scene(10)
    scene = dev.upload_scene(bs.lower, bs.upper, bs.matrices)
    plan = dev.batched_plan(bs)
    masks = dev.DeviceArray(bs.mask_words, np.uint64)
    vps = dev.DeviceArray.from_host(bs.vps.reshape(-1, 16))
    rc = lib.dpc_cull_batch(plan.handle, vps.ptr, masks.ptr, bs.mask_words, None)
    assert rc == 6 and b"bound" in lib.dpc_last_error()
    short = dev.scene_desc(scene)
    short.n_objects = scene.n_objects - 1
    table = (_lib.SceneDesc * 1)(short)
    assert lib.dpc_batch_plan_bind(plan.handle, table, 1) == 1 and b"reaches object" in lib.dpc_last_error()
    plan.bind([scene])
    rc = lib.dpc_cull_batch(plan.handle, vps.ptr, masks.ptr, bs.mask_words - 1, None)
    assert rc == 1 and b"mask buffer" in lib.dpc_last_error()
    plan.destroy()
    for d in (masks, vps):
        d.free()
    scene.free()


def test_groups_across_scenes(orc):
    """One launch over groups that live in different device scenes (each with its
    own arrays), as the serving loop batches requests of resident scenes."""
    a = small_scene(11)
    b = small_scene(12)
    sa = dev.upload_scene(a.lower, a.upper, a.matrices)
    sb = dev.upload_scene(b.lower, b.upper, b.matrices)
    vps = np.concatenate([a.vps, b.vps])
    fa = a.vps.shape[0]
    groups = [(0, f, n) for f, n in zip(a.group_first, a.group_count)] + \
             [(1, f, n) for f, n in zip(b.group_first, b.group_count)]
    ga = len(a.group_first)
    pairs = list(zip(a.pair_frustum, a.pair_group, a.pair_offset)) + \
        [(f + fa, g + ga, o + a.mask_words) for f, g, o in zip(b.pair_frustum, b.pair_group, b.pair_offset)]
    plan = dev.BatchPlan(groups, pairs, vps.shape[0])
    plan.bind([sa, sb])
    d_vps = dev.DeviceArray.from_host(vps.reshape(-1, 16))
    masks = dev.DeviceArray(plan.mask_words, np.uint64)
    plan.cull(d_vps, masks)
    got = masks.download()
    want = np.concatenate([oracle_pairs(orc, a), oracle_pairs(orc, b)])
    assert mismatches(got, want, plan.mask_words * 64) == 0
    plan.destroy()
    for d in (d_vps, masks):
        d.free()
    sa.free()
    sb.free()


def test_plan_rejects_bad_pairs():
    lib = _lib.load()
    groups = (_lib.GroupRange * 1)(_lib.GroupRange(0, 10))
    pairs = (_lib.CullPair * 1)(_lib.CullPair(3, 0, 0))
    plan = C.c_void_p()
    assert lib.dpc_batch_plan_create(groups, 1, pairs, 1, 2, None, C.byref(plan)) == 1
    bad = _lib.BatchConfig(3, 256, 8, 0)
    pairs = (_lib.CullPair * 1)(_lib.CullPair(0, 0, 0))
    assert lib.dpc_batch_plan_create(groups, 1, pairs, 1, 1, C.byref(bad), C.byref(plan)) == 1


def test_valu_probe_rates_are_plausible():
    """dpc_probe_valu_rate (the calibration the batched cull's issue fraction is
    read against): independent FP32 adds and multiplies issue at a positive rate
    no higher than 64 lanes per CU per clock at the device's clock."""
    import ctypes as C
    from pipeline_amd import device as dev
    lib = _lib.load()
    info = dev.device_info(0)
    ceiling = info["cu_count"] * 64 * info["clock_khz"] * 1e3
    for op in (0, 1):
        r = C.c_double()
        _lib.check(lib.dpc_probe_valu_rate(op, 2000, C.byref(r), None))
        assert 0.25 * ceiling < r.value <= 1.02 * ceiling, (op, r.value, ceiling)
    assert lib.dpc_probe_valu_rate(2, 10, C.byref(C.c_double()), None) == 1  # DPC_ERR_INVALID_ARGUMENT

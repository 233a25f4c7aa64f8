"""GPU parity of single-frustum culling (dpc_cull) with the CPU oracle of
dp::culling::cpu::ManagerImpl::cull.  Bit-exact: 0 mismatching visibility bits."""
import numpy as np
import pytest

from gpu_util import bits_of, gpu_cull_flat, mismatches
from pipeline_amd import device as dev
from pipeline_amd import scenes

pytestmark = pytest.mark.gpu
F = np.float32
ID = np.eye(4, dtype=F)


def oracle_mask(orc, lower, upper, matrices, vp, xform=None):
    l4, e4 = scenes.extents(lower, upper)
    x = np.arange(lower.shape[0]) if xform is None else xform
    return orc.cull_group(l4, e4, x, matrices, vp)


def test_flat_10k(orc):
    s = scenes.flat_scene_10k()
    got = gpu_cull_flat(s.lower, s.upper, s.matrices, s.vp)
    want = oracle_mask(orc, s.lower, s.upper, s.matrices, s.vp)
    assert mismatches(got, want, s.n) == 0
    assert 0 < bits_of(want, s.n).sum() < s.n


def test_flat_1m(orc, cull_shape):
    s = scenes.flat_scene_1m()
    got = gpu_cull_flat(s.lower, s.upper, s.matrices, s.vp)
    want = oracle_mask(orc, s.lower, s.upper, s.matrices, s.vp)
    n_mis = mismatches(got, want, s.n)
    print(f"flat 1M: {s.n} objects, {int(bits_of(want, s.n).sum())} visible, {n_mis} mismatching bits")
    assert n_mis == 0


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 4097])
def test_ragged_sizes(orc, n, cull_shape):
    s = scenes.flat_scene(n, 100 + n, (-30.0, 30.0), 70.0, 1.3, 0.1, 60.0)
    got = gpu_cull_flat(s.lower, s.upper, s.matrices, s.vp)
    assert mismatches(got, oracle_mask(orc, s.lower, s.upper, s.matrices, s.vp), n) == 0


def test_shared_matrices_through_indices(orc, cull_shape):
    """Objects reference matrices by transform index (many objects per matrix)."""
    rng = np.random.default_rng(5)
    n, nm = 50_000, 1_000
    lower, upper = scenes.random_boxes(rng, n, (-20.0, 20.0))
    mats = scenes.trs_matrices(rng, nm)
    xform = rng.integers(0, nm, n).astype(np.uint32)
    vp = scenes.perspective_vp([0, 0, 30], [0, 0, -1], 60, 1.5, 0.1, 100)
    got = gpu_cull_flat(lower, upper, mats, vp, xform)
    want = oracle_mask(orc, lower, upper, mats, vp, xform)
    assert mismatches(got, want, n) == 0


def test_out_of_range_index_stays_visible(orc, cull_shape):
    """A transform index beyond the matrix array is never read; the object is reported visible."""
    lower = np.array([[10, 10, 10], [10, 10, 10]], F)
    upper = lower + 1
    mats = ID[None].copy()
    got = gpu_cull_flat(lower, upper, mats, ID, np.array([0, 7], np.uint32))
    assert bits_of(got, 2).tolist() == [0, 1]


def _known_cases():
    """The hand-derived cases of test_oracle_known_answers, run on the device."""
    e = F(3 * 2.0 ** -24)
    s, t = F(1 + 2.0 ** -23), F(1 - 2.0 ** -24)
    cases = [
        # (lower, upper, M, VP, expected visible)
        ((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), ID, ID, True),
        ((2, -0.5, -0.5), (3, 0.5, 0.5), ID, ID, False),
        ((1, -0.5, -0.5), (2, 0.5, 0.5), ID, ID, False),          # <= on the plane
        ((-5, -5, -5), (5, 5, 5), ID, ID, True),
        ((0.5, 0.5, 0.5), (2, 2, 2), ID, np.diag([1, 1, 1, -1]).astype(F), True),   # else-if
        ((-3, 0, 0), (-1, 0.5, 1), np.array([[-1, 0, 0, 0], [0, 1, 0, 0], [-e, 0, 1, 0], [0, 0, 0, 1]], F),
         np.diag([1, 1, 1, F(1) - e]).astype(F), False),          # corner add tree
        ((3, 0, 0), (3, 0, 0), np.diag([s, s, s, 1]).astype(F),
         np.diag([t, 1, 1, F(3 + 2.0 ** -22)]).astype(F), False),  # two-stage rounding
    ]
    m_nan = ID.copy()
    m_nan[3, 3] = np.nan
    cases.append(((10, 10, 10), (11, 11, 11), m_nan, ID, True))   # NaN w
    return cases


@pytest.mark.parametrize("case", range(8))
def test_known_answer_cases_on_device(orc, case):
    lower, upper, m, vp, want = _known_cases()[case]
    lo = np.array([lower], F)
    up = np.array([upper], F)
    got = gpu_cull_flat(lo, up, m[None], vp)
    assert bool(bits_of(got, 1)[0]) is want
    assert mismatches(got, oracle_mask(orc, lo, up, m[None], vp), 1) == 0


def test_special_values(orc, cull_shape):
    """Infinities, NaNs, denormals and signed zeros in boxes and matrices: the
    device evaluates them exactly as the oracle does (no flush-to-zero)."""
    rng = np.random.default_rng(9)
    n = 4096
    lower, upper = scenes.random_boxes(rng, n, (-3.0, 3.0), (0.01, 1.0))
    mats = scenes.trs_matrices(rng, n)
    specials = np.array([np.inf, -np.inf, np.nan, 1e-40, -1e-40, 0.0, -0.0, 3e38, -3e38], F)
    pick = rng.random(mats.shape) < 0.01
    mats[pick] = rng.choice(specials, pick.sum())
    pick = rng.random(lower.shape) < 0.01
    lower[pick] = rng.choice(specials, pick.sum())
    vp = scenes.perspective_vp([0, 0, 5], [0, 0, -1], 60, 1.0, 0.1, 20)
    vp[1, 1] = F(1e-41)   # a denormal in the projection
    got = gpu_cull_flat(lower, upper, mats, vp)
    assert mismatches(got, oracle_mask(orc, lower, upper, mats, vp), n) == 0


def test_repeatable(orc):
    s = scenes.flat_scene(100_000, 31, (-100.0, 100.0))
    a = gpu_cull_flat(s.lower, s.upper, s.matrices, s.vp)
    b = gpu_cull_flat(s.lower, s.upper, s.matrices, s.vp)
    assert np.array_equal(a, b)


def test_layout_round_trip_and_scatter():
    """dpc_matrices_from_aos / _to_aos / _scatter and dpc_boxes_from_lower_upper."""
    import ctypes as C
    from pipeline_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(4)
    n = 1000
    lower, upper = scenes.random_boxes(rng, n)
    mats = scenes.trs_matrices(rng, n)
    scene = dev.upload_scene(lower, upper, mats)
    planes = scene.planes.download()
    assert np.array_equal(planes, dev.matrices_to_planes(mats))
    a, b = scene.box_a.download(), scene.box_b.download()
    ext = (upper - lower).astype(F)
    assert np.array_equal(a[:, :3], lower) and np.array_equal(a[:, 3], ext[:, 0])
    assert np.array_equal(b, ext[:, 1:])
    # scatter a delta, then read back through a slot map
    idx = rng.choice(n, 100, replace=False).astype(np.uint32)
    new = scenes.trs_matrices(rng, 100)
    d_idx = dev.DeviceArray.from_host(idx)
    d_new = dev.DeviceArray.from_host(new.reshape(-1, 16))
    _lib.check(lib.dpc_matrices_scatter(scene.planes.ptr, scene.ld, d_idx.ptr, d_new.ptr, 100, None))
    out = dev.DeviceArray((100, 16), F)
    _lib.check(lib.dpc_matrices_to_aos(scene.planes.ptr, scene.ld, d_idx.ptr, 100, out.ptr, 64, None))
    assert np.array_equal(out.download().reshape(100, 4, 4), new)
    full = dev.DeviceArray((n, 16), F)
    _lib.check(lib.dpc_matrices_to_aos(scene.planes.ptr, scene.ld, None, n, full.ptr, 64, None))
    want = mats.copy()
    want[idx] = new
    assert np.array_equal(full.download().reshape(n, 4, 4), want)
    for d in (d_idx, d_new, out, full):
        d.free()
    scene.free()


def test_invalid_arguments_are_errors(lib):
    import ctypes as C
    from pipeline_amd import _lib
    vp = dev.vp_arg(ID)
    rc = lib.dpc_cull(None, None, None, 10, None, 10, vp, None, None)
    assert rc == 1 and b"null" in lib.dpc_last_error()
    d = dev.DeviceArray(64, F)
    rc = lib.dpc_cull(d.addr + 4, d.ptr, None, 1, d.ptr, 1, vp, d.ptr, None)   # misaligned box_a
    assert rc == 1
    rc = lib.dpc_cull(d.ptr, d.ptr, None, 100, d.ptr, 10, vp, d.ptr, None)      # ld < n without indices
    assert rc == 1
    assert lib.dpc_cull(d.ptr, d.ptr, None, 0, d.ptr, 0, vp, d.ptr, None) == 0  # empty group: no launch
    d.free()

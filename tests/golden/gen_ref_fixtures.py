"""Writes the golden vectors in tests/golden/ from the reference's own compiled
code (oracle/_ref/libdpref.so: dp/math operators and camera builders,
dp::util::BitArray, dp::transform::Tree, compiled from /root/reference by
oracle/Makefile).

Run in the container that has the reference sources:
    make -C oracle ref && python tests/golden/gen_ref_fixtures.py [math bitarray tree invert]

The fixtures are data only (inputs and the reference's outputs); the reference
itself is not needed to check against them.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402

OUT = Path(__file__).resolve().parent


def _p(a):
    return a.ctypes.data


def wide_floats(rng, shape):
    """Floats over a wide exponent range with random signs, plus exact small
    integers and zeros, so that rounding differences between evaluation orders show."""
    x = rng.uniform(-1, 1, shape) * np.exp2(rng.integers(-20, 21, shape))
    pick = rng.random(shape)
    x = np.where(pick < 0.05, 0.0, x)
    x = np.where((pick >= 0.05) & (pick < 0.1), np.round(x), x)
    return x.astype(np.float32)


def gen_math(ref, rng):
    n = 2000
    v = wide_floats(rng, (n, 4))
    m = wide_floats(rng, (n, 16))
    a = wide_floats(rng, (n, 16))
    s = wide_floats(rng, (n,))
    vm = np.empty((n, 4), np.float32)
    mm = np.empty((n, 16), np.float32)
    sv = np.empty((n, 4), np.float32)
    va = np.empty((n, 4), np.float32)
    for i in range(n):
        ref.ref_vec4_mul_mat44(_p(v[i]), _p(m[i]), _p(vm[i]))
        ref.ref_mat44_mul(_p(a[i]), _p(m[i]), _p(mm[i]))
        ref.ref_scalar_mul_vec4(float(s[i]), _p(v[i]), _p(sv[i]))
        ref.ref_vec4_add(_p(v[i]), _p(vm[i]), _p(va[i]))
    # camera builders
    persp_args = np.array([[60, 16 / 9, 0.1, 500], [45, 16 / 9, 0.1, 2000], [90, 1, 0.1, 2000],
                           [30, 1.5, 0.5, 100], [75, 0.75, 1, 3000]], np.float32)
    persp = np.empty((len(persp_args), 16), np.float32)
    for i, (f, asp, zn, zf) in enumerate(persp_args):
        ref.ref_make_perspective(float(f), float(asp), float(zn), float(zf), _p(persp[i]))
    ortho_args = np.array([[-50, 50, -50, 50, 0.1, 4000], [-10, 30, -5, 7, 1, 100]], np.float32)
    ortho = np.empty((len(ortho_args), 16), np.float32)
    for i, a6 in enumerate(ortho_args):
        ref.ref_make_ortho(*[float(x) for x in a6], _p(ortho[i]))
    look_args = np.array([[0, 0, 0, 0, 0, -1, 0, 1, 0], [10, -20, 5, 0, 0, 0, 0, 1, 0],
                          [-3, 4, 40, 1, 2, 3, 0, 0, 1]], np.float32)
    look = np.empty((len(look_args), 16), np.float32)
    for i, a9 in enumerate(look_args):
        eye, center, up = a9[0:3].copy(), a9[3:6].copy(), a9[6:9].copy()  # keep the buffers alive for the call
        ref.ref_make_lookat(_p(eye), _p(center), _p(up), _p(look[i]))
    q = rng.normal(size=(50, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    tr = rng.uniform(-10, 10, (50, 3)).astype(np.float32)
    sc = np.repeat(rng.uniform(0.5, 2, (50, 1)), 3, axis=1).astype(np.float32)
    qm = np.empty((50, 16), np.float32)
    for i in range(50):
        ref.ref_set_mat_quat(_p(q[i]), _p(tr[i]), _p(sc[i]), _p(qm[i]))
    np.savez_compressed(OUT / "ref_math.npz", v=v, m=m, a=a, s=s, vec_mat=vm, mat_mat=mm, scalar_vec=sv,
                        vec_add=va, persp_args=persp_args, persp=persp, ortho_args=ortho_args, ortho=ortho,
                        look_args=look_args, look=look, quat=q, quat_trans=tr, quat_scale=sc, quat_mat=qm)


def gen_bitarray(ref, rng):
    sizes = [1, 5, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 1000]
    srcs, setbits, clz, trav_idx, trav_cnt = [], [], [], [], []
    for n in sizes:
        words32 = 2 * ((n + 63) // 64)
        for density in (0.0, 0.05, 0.5, 1.0):
            src = np.zeros(words32, np.uint32)
            bits = (rng.random(words32 * 32) < density).astype(np.uint8)
            # garbage above n in the source: setBits must clear it
            src[:] = np.packbits(bits, bitorder="little").view(np.uint32)[:words32]
            out = np.zeros((n + 63) // 64, np.uint64)
            ref.ref_bitarray_setbits(_p(src), n, _p(out))
            c = ref.ref_bitarray_count_leading_zeroes(_p(src), n)
            idx = np.zeros(max(n, 1), np.uint32)
            k = ref.ref_bitarray_traverse(_p(src), n, _p(idx))
            srcs.append(src)
            setbits.append(out)
            clz.append(c)
            trav_idx.append(idx[:k])
            trav_cnt.append(k)
    meta = np.array([[n, d] for n in sizes for d in range(4)], np.int64)
    np.savez_compressed(OUT / "ref_bitarray.npz", meta=meta, src=np.concatenate(srcs),
                        src_len=np.array([len(s) for s in srcs]), setbits=np.concatenate(setbits),
                        setbits_len=np.array([len(s) for s in setbits]), clz=np.array(clz, np.int64),
                        trav=np.concatenate(trav_idx), trav_len=np.array(trav_cnt, np.int64))


def random_local(rng):
    from pipeline_amd.scenes import trs_matrices
    return trs_matrices(rng, 1, scale=(0.8, 1.25), trans=(-10, 10))[0].reshape(16)


def gen_tree(ref, rng):
    """A trace of Tree operations: adds (depth-first and random), local updates,
    removals of leaves and of inner nodes (orphaning subtrees), re-adds that reuse
    freed indices, and a compute after each phase."""
    t = ref.ref_tree_create()
    ops = []           # (kind, a, b): kind 0=add(parent=a -> index b), 1=update(a), 2=remove(a), 3=compute
    mats = []          # matrix per op (zeros for remove/compute)
    live = [0]
    computes = []

    def add(parent):
        m = random_local(rng)
        idx = C.c_uint32()
        rc = ref.ref_tree_add(t, parent, _p(m), C.byref(idx))
        ops.append((0, parent, idx.value if rc == 0 else 0xFFFFFFFF))
        mats.append(m)
        if rc == 0:
            live.append(idx.value)
        return rc

    def compute():
        count = ref.ref_tree_count(t)
        event = np.zeros(count, np.uint8)
        ref.ref_tree_compute(t, _p(event))
        world = np.empty((count, 16), np.float32)
        ref.ref_tree_world(t, _p(world))
        nl = ref.ref_tree_level_count(t)
        levels = []
        for l in range(nl):
            k = ref.ref_tree_level_size(t, l)
            e = np.zeros(2 * max(k, 1), np.uint32)
            ref.ref_tree_level_entries(t, l, _p(e))
            levels.append(e[: 2 * k])
        hi = max(live) + 1
        computes.append({"world": world[:hi].copy(), "event": event[:hi].copy(), "levels": levels})
        ops.append((3, len(computes) - 1, 0))
        mats.append(np.zeros(16, np.float32))

    # phase 1: depth-first build (what a scene traversal does)
    def dfs(parent, depth):
        for _ in range(int(rng.integers(1, 4)) if depth < 5 else 0):
            add(parent)
            dfs(live[-1], depth + 1)
    dfs(0, 0)
    compute()
    # phase 2: random adds under random live nodes
    for _ in range(60):
        add(int(rng.choice(live)))
    compute()
    # phase 3: update some locals
    for idx in rng.choice(live[1:], 25, replace=False):
        m = random_local(rng)
        ref.ref_tree_update_local(t, int(idx), _p(m))
        ops.append((1, int(idx), 0))
        mats.append(m)
    compute()
    # phase 4: no change (nothing dirty)
    compute()
    # phase 5: removals (inner nodes orphan their subtrees) and re-adds reusing indices
    for idx in rng.choice(live[1:], 15, replace=False):
        rc = ref.ref_tree_remove(t, int(idx))
        ops.append((2, int(idx), rc & 0xFFFFFFFF))
        mats.append(np.zeros(16, np.float32))
        if rc == 0:
            live.remove(int(idx))
    for _ in range(20):
        add(int(rng.choice(live)))
    # invalid operations: parent that does not exist, removing a freed index
    add(60000)
    rc = ref.ref_tree_remove(t, 60001)
    ops.append((2, 60001, rc & 0xFFFFFFFF))
    mats.append(np.zeros(16, np.float32))
    for idx in rng.choice(live[1:], 10, replace=False):
        m = random_local(rng)
        ref.ref_tree_update_local(t, int(idx), _p(m))
        ops.append((1, int(idx), 0))
        mats.append(m)
    compute()
    ref.ref_tree_destroy(t)

    arrays = {"ops": np.array(ops, np.uint32), "op_mats": np.stack(mats).astype(np.float32)}
    for k, c in enumerate(computes):
        arrays[f"c{k}_world"] = c["world"]
        arrays[f"c{k}_event"] = c["event"]
        lens = np.array([len(l) // 2 for l in c["levels"]], np.int64)
        arrays[f"c{k}_level_len"] = lens
        arrays[f"c{k}_levels"] = (np.concatenate(c["levels"]) if c["levels"] else np.zeros(0, np.uint32))
    arrays["n_computes"] = np.array(len(computes))
    np.savez_compressed(OUT / "ref_tree.npz", **arrays)
    write_tree_trace_bin(arrays)


def write_tree_trace_bin(arrays):
    """The same trace as a flat little-endian file for the C++ tests
    (tests/cpp/tree_trace.hpp reads it):
      "DPTT" u32 version=1
      u32 n_ops, then n_ops x (u32 kind, u32 a, u32 b, 16 x f32 matrix)
      u32 n_computes, then per compute:
        u32 n_levels, n_levels x u32 size, entries (2 x u32 each),
        u32 n_world, n_world x 16 f32 world matrices, n_world x u8 event bytes"""
    import struct
    ops, mats = arrays["ops"], arrays["op_mats"]
    out = bytearray(b"DPTT") + struct.pack("<I", 1)
    out += struct.pack("<I", len(ops))
    for (kind, a, b), m in zip(ops, mats):
        out += struct.pack("<III", int(kind), int(a), int(b)) + np.asarray(m, "<f4").tobytes()
    n_c = int(arrays["n_computes"])
    out += struct.pack("<I", n_c)
    for k in range(n_c):
        lens = arrays[f"c{k}_level_len"]
        out += struct.pack("<I", len(lens)) + np.asarray(lens, "<u4").tobytes()
        out += np.asarray(arrays[f"c{k}_levels"], "<u4").tobytes()
        w = arrays[f"c{k}_world"]
        out += struct.pack("<I", w.shape[0]) + np.asarray(w, "<f4").tobytes()
        out += np.asarray(arrays[f"c{k}_event"], np.uint8).tobytes()
    (OUT / "ref_tree_trace.bin").write_bytes(bytes(out))


def gen_invert(ref, rng):
    """dp::math::invertTranspose on world-like matrices (products of rotation x
    scale x translation chains), on matrices with entries over a wide exponent
    range, and on singular, non-finite and underflowing cases."""
    from pipeline_amd import scenes
    cases = []
    # chains of 1-6 local TRS matrices, multiplied by the reference's Mat44 product
    for depth in range(1, 7):
        loc = scenes.trs_matrices(rng, 300 * depth, scale=(0.8, 1.25), trans=(-10, 10)).reshape(-1, depth, 16)
        for chain in loc:
            w = chain[0].copy()
            for k in range(1, depth):
                nxt = np.empty(16, np.float32)
                ref.ref_mat44_mul(_p(chain[k]), _p(w), _p(nxt))
                w = nxt
            cases.append(w)
    cases += list(wide_floats(rng, (1500, 16)))
    special = []
    base = wide_floats(rng, (40, 16)).reshape(-1, 4, 4)
    for b in base[:10]:  # two equal rows: every 2x2 minor of rows 0,1 is exactly 0
        b = b.copy()
        b[1] = b[0]
        special.append(b.reshape(16))
    for b in base[10:20]:  # a zero column
        b = b.copy()
        b[:, 2] = 0
        special.append(b.reshape(16))
    special.append(np.zeros(16, np.float32))
    for b in base[20:30]:  # a NaN or an infinity somewhere
        b = b.copy().reshape(16)
        b[int(rng.integers(16))] = np.nan if rng.random() < 0.5 else np.inf
        special.append(b)
    for sc in (1e-10, 1e-12, 1e-20, 1e-30, 1e15, 1e20):  # det underflows or overflows
        special.append((np.eye(4, dtype=np.float64) * sc).astype(np.float32).reshape(16))
        m = scenes.trs_matrices(rng, 1)[0].astype(np.float64)
        m[:3, :3] *= sc
        special.append(m.astype(np.float32).reshape(16))
    cases += special
    m = np.ascontiguousarray(np.stack(cases), np.float32)
    out = np.full(m.shape, np.nan, np.float32)
    ok = np.zeros(m.shape[0], np.uint8)
    for i in range(m.shape[0]):
        ok[i] = ref.ref_invert_transpose(_p(m[i]), _p(out[i]))
    np.savez_compressed(OUT / "ref_invert.npz", m=m, ok=ok, out=out)


GENERATORS = {"math": gen_math, "bitarray": gen_bitarray, "tree": gen_tree, "invert": gen_invert}


def main():
    ref = orc.load_ref()
    which = sys.argv[1:] or list(GENERATORS)
    rng = np.random.default_rng(20261202)
    for name, gen in GENERATORS.items():
        # every set draws from its own stream, so regenerating one leaves the others as they are
        if name in which:
            gen(ref, np.random.default_rng([20261202, list(GENERATORS).index(name)]) if name == "invert" else rng)
    for f in sorted(OUT.glob("ref_*.npz")):
        print(f.name, f.stat().st_size, "bytes")


if __name__ == "__main__":
    main()

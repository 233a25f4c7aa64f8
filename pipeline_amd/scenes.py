"""Seeded synthetic workloads of the scene-update path.

All inputs are FP32.  Matrices are 4x4 row-major in the row-vector convention
of dp::math (v' = v * M).  The camera builders follow dp/math/Matmnt.h
(makeLookAt :1250-1280, makeOrtho :1297-1306, makeFrustum :1325-1345,
makePerspective :1358-1368) and are evaluated in float64, then rounded once to
float32; the resulting matrices are the shared input of the CPU oracle and the
GPU, byte for byte.

Workloads
  flat_scene        N objects, each with its own world matrix (rotation x uniform
                    scale x translation), box centres uniform in a cube
  tree_scene        random transform tree (level-major) with objects on its leaves
  batched_scene     objects in groups of log-uniform size, frusta of three kinds,
                    each frustum paired with 1-8 groups
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------------------
# camera builders (float64, rounded to float32 at the end)
# ---------------------------------------------------------------------------
def _normalize(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def make_frustum(l, r, b, t, n, f):
    v0 = (r + l) / (r - l)
    v1 = (t + b) / (t - b)
    v2 = -(f + n) / (f - n)
    v3 = -2.0 * f * n / (f - n)
    v4 = 2.0 * n / (r - l)
    v5 = 2.0 * n / (t - b)
    return np.array([[v4, 0, 0, 0], [0, v5, 0, 0], [v0, v1, v2, -1], [0, 0, v3, 0]], np.float64)


def make_perspective(fovy_deg, aspect, znear, zfar):
    tanfov = np.tan(np.deg2rad(fovy_deg) * 0.5)
    r = tanfov * aspect * znear
    t = tanfov * znear
    return make_frustum(-r, r, -t, t, znear, zfar)


def make_ortho(l, r, b, t, n, f):
    return np.array([[2 / (r - l), 0, 0, 0], [0, 2 / (t - b), 0, 0], [0, 0, -2 / (f - n), 0],
                     [-(r + l) / (r - l), -(t + b) / (t - b), -(f + n) / (f - n), 1]], np.float64)


def make_lookat(eye, center, up):
    eye = np.asarray(eye, np.float64)
    f = _normalize(np.asarray(center, np.float64) - eye)
    s = _normalize(np.cross(f, up))
    u = np.cross(s, f)
    trans = np.eye(4)
    trans[3, :3] = -eye
    ori = np.eye(4)
    ori[0, :3] = [s[0], u[0], -f[0]]
    ori[1, :3] = [s[1], u[1], -f[1]]
    ori[2, :3] = [s[2], u[2], -f[2]]
    return trans @ ori


def view_projection(view, proj) -> np.ndarray:
    """world -> clip in the row-vector convention: VP = View * Projection."""
    return (np.asarray(view, np.float64) @ np.asarray(proj, np.float64)).astype(np.float32)


def _up_for(direction):
    d = _normalize(direction)
    return np.array([0.0, 0.0, 1.0]) if abs(d[1]) > 0.95 else np.array([0.0, 1.0, 0.0])


def perspective_vp(eye, direction, fovy, aspect, znear, zfar) -> np.ndarray:
    eye = np.asarray(eye, np.float64)
    return view_projection(make_lookat(eye, eye + _normalize(direction), _up_for(direction)),
                           make_perspective(fovy, aspect, znear, zfar))


# ---------------------------------------------------------------------------
# object matrices and boxes
# ---------------------------------------------------------------------------
def random_unit_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotations (Shoemake 1992), rows (x, y, z, w)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)


def quat_rows(q: np.ndarray) -> np.ndarray:
    """Rotation rows of setMat(Mat33, Quat) (dp/math/Matmnt.h:1100-1114), (n, 3, 3)."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.empty((q.shape[0], 3, 3))
    r[:, 0] = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)], 1)
    r[:, 1] = np.stack([2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)], 1)
    r[:, 2] = np.stack([2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)], 1)
    return r


def trs_matrices(rng: np.random.Generator, n: int, scale=(0.5, 2.0), trans=(-5.0, 5.0)) -> np.ndarray:
    """setMat(Mat44, quat, translation, scale) (Matmnt.h:527-537) with a uniform scale:
    rows 0..2 = s * R[k] (w = 0), row 3 = (t, 1).  Returns (n, 4, 4) float32."""
    rot = quat_rows(random_unit_quaternions(rng, n))
    s = rng.uniform(scale[0], scale[1], n)
    t = rng.uniform(trans[0], trans[1], (n, 3))
    m = np.zeros((n, 4, 4))
    m[:, :3, :3] = rot * s[:, None, None]
    m[:, 3, :3] = t
    m[:, 3, 3] = 1.0
    return m.astype(np.float32)


def random_boxes(rng: np.random.Generator, n: int, center=(-100.0, 100.0), half=(0.1, 2.0)):
    """lower/upper (n, 3) float32 from centres and half-extents; lower = c - h and
    upper = c + h are rounded to float32 before the extent is formed."""
    c = rng.uniform(center[0], center[1], (n, 3)).astype(np.float32)
    h = rng.uniform(half[0], half[1], (n, 3)).astype(np.float32)
    return (c - h).astype(np.float32), (c + h).astype(np.float32)


def extents(lower: np.ndarray, upper: np.ndarray):
    """(lowerLeft4, extent4) as ManagerBitSet::objectSetBoundingBox stores them
    (dp/culling/src/ManagerBitSet.cpp:99-109): (lower, 1) and (upper - lower, 0)."""
    n = lower.shape[0]
    l4 = np.empty((n, 4), np.float32)
    e4 = np.empty((n, 4), np.float32)
    l4[:, :3] = lower
    l4[:, 3] = 1.0
    e4[:, :3] = (upper.astype(np.float32) - lower.astype(np.float32)).astype(np.float32)
    e4[:, 3] = 0.0
    return l4, e4


# ---------------------------------------------------------------------------
# flat scenes (one matrix per object)
# ---------------------------------------------------------------------------
@dataclass
class FlatScene:
    lower: np.ndarray      # (n, 3) f32
    upper: np.ndarray      # (n, 3) f32
    matrices: np.ndarray   # (n, 4, 4) f32, object i uses matrix i
    vp: np.ndarray         # (4, 4) f32
    name: str = ""

    @property
    def n(self) -> int:
        return self.lower.shape[0]


def flat_scene(n: int, seed: int, center=(-100.0, 100.0), fovy=60.0, aspect=16 / 9, znear=0.1, zfar=500.0,
               eye=None, direction=None, name="") -> FlatScene:
    rng = np.random.default_rng(seed)
    lower, upper = random_boxes(rng, n, center)
    mats = trs_matrices(rng, n)
    if eye is None:
        eye = np.zeros(3)
    if direction is None:
        direction = np.array([0.0, 0.0, -1.0])
    vp = perspective_vp(eye, direction, fovy, aspect, znear, zfar)
    return FlatScene(lower, upper, mats, vp, name)


def flat_scene_10k(seed: int = 0) -> FlatScene:
    """10,000 objects, centres U[-100,100]^3, one perspective frustum (fov 60, 16:9,
    near 0.1, far 500) at the origin looking down -z."""
    return flat_scene(10_000, seed, (-100.0, 100.0), 60.0, 16 / 9, 0.1, 500.0, name="flat-10k")


def flat_scene_1m(seed: int = 1, n: int = 1 << 20) -> FlatScene:
    """1,048,576 objects, centres U[-1000,1000]^3, fov 45, far 2000, random eye in
    [-50,50]^3 and random look direction."""
    rng = np.random.default_rng(seed + 1_000_003)
    eye = rng.uniform(-50, 50, 3)
    d = rng.normal(size=3)
    return flat_scene(n, seed, (-1000.0, 1000.0), 45.0, 16 / 9, 0.1, 2000.0, eye, d,
                      name=f"flat-{n}")


# ---------------------------------------------------------------------------
# transform tree
# ---------------------------------------------------------------------------
@dataclass
class TreeScene:
    """A transform tree in level-major slot order plus the reference's view of it.

    Slot 0 is the virtual root (identity, level -1).  Slots 1.. hold the nodes
    level by level; within a level, nodes are in the reference's list order.
    The reference adds transforms in depth-first pre-order (as a scene-graph
    traversal does), so transform index = 1 + pre-order rank."""
    parent_slot: np.ndarray     # (n_slots,) u32, parent slot of each slot (0 for the root itself)
    level_offsets: np.ndarray   # (n_levels + 1,) u64 slot ranges
    local: np.ndarray           # (n_slots, 4, 4) f32 local matrices by slot (slot 0 = identity)
    index_of_slot: np.ndarray   # (n_slots,) u32 transform index of each slot
    obj_lower: np.ndarray       # (m, 3)
    obj_upper: np.ndarray       # (m, 3)
    obj_slot: np.ndarray        # (m,) u32 slot of each object's transform
    vp: np.ndarray

    @property
    def n_slots(self) -> int:
        return self.parent_slot.shape[0]

    def reference_levels(self):
        """(level_offsets, entries) with entries = (parent index, transform index) pairs in
        the reference's per-level order, as dp::transform::Tree keeps them."""
        entries = np.stack([self.index_of_slot[self.parent_slot], self.index_of_slot], axis=1)[1:]
        return self.level_offsets - 1, np.ascontiguousarray(entries, np.uint32)

    def local_by_index(self) -> np.ndarray:
        out = np.zeros_like(self.local)
        out[self.index_of_slot] = self.local
        return out


def tree_scene(n_nodes: int, seed: int, max_depth: int = 16, p: float = 0.4, max_children: int = 8,
               objects_per_leaf: int = 4, half=(0.05, 1.0), roots: int = 8) -> TreeScene:
    rng = np.random.default_rng(seed)
    # Level by level (breadth first): node counts grow by ~2.45 per level.  Level 0
    # holds `roots` top-level transforms under the virtual root.
    roots = max(1, min(roots, n_nodes))
    level_parents = [np.zeros(roots, np.int64)]
    total = roots
    while total < n_nodes and len(level_parents) < max_depth:
        prev = level_parents[-1].shape[0]
        k = np.minimum(rng.geometric(p, prev), max_children)
        parents = np.repeat(np.arange(prev), k)
        if total + parents.shape[0] > n_nodes:
            parents = parents[: n_nodes - total]
        if parents.shape[0] == 0:
            break
        level_parents.append(parents)
        total += parents.shape[0]
    sizes = np.array([lp.shape[0] for lp in level_parents], np.int64)
    level_start = np.concatenate([[1], 1 + np.cumsum(sizes)])  # slot of each level's first node
    n_slots = int(level_start[-1])
    parent_slot = np.zeros(n_slots, np.uint32)  # level 0 (and slot 0 itself) -> virtual root
    for l in range(1, len(level_parents)):
        parent_slot[level_start[l]:level_start[l + 1]] = level_start[l - 1] + level_parents[l]

    # Depth-first pre-order rank of every slot: subtree sizes bottom-up, then
    # rank(child) = rank(parent) + 1 + sizes of the earlier siblings.
    sub = np.ones(n_slots, np.int64)
    sub[0] = 0
    for l in range(len(level_parents) - 1, 0, -1):
        s0, s1 = level_start[l], level_start[l + 1]
        np.add.at(sub, parent_slot[s0:s1], sub[s0:s1])
    rank = np.zeros(n_slots, np.int64)
    rank[1:1 + roots] = np.concatenate([[0], np.cumsum(sub[1:roots])])  # top-level subtrees in order
    for l in range(1, len(level_parents)):
        s0, s1 = level_start[l], level_start[l + 1]
        ps = parent_slot[s0:s1].astype(np.int64)
        cs = np.cumsum(sub[s0:s1])
        excl = cs - sub[s0:s1]
        # children of one parent are contiguous: subtract the running sum at the group start
        first = np.r_[True, ps[1:] != ps[:-1]]
        group_base = np.maximum.accumulate(np.where(first, excl, 0))
        rank[s0:s1] = rank[ps] + 1 + (excl - group_base)
    index_of_slot = np.zeros(n_slots, np.uint32)
    index_of_slot[1:] = (1 + rank[1:]).astype(np.uint32)

    local = np.zeros((n_slots, 4, 4), np.float32)
    local[0] = np.eye(4, dtype=np.float32)
    local[1:] = trs_matrices(rng, n_slots - 1, scale=(0.8, 1.25), trans=(-10.0, 10.0))

    has_child = np.zeros(n_slots, bool)
    has_child[parent_slot[1:]] = True
    has_child[0] = True
    leaves = np.nonzero(~has_child)[0].astype(np.uint32)
    obj_slot = np.repeat(leaves, objects_per_leaf)
    m = obj_slot.shape[0]
    lower, upper = random_boxes(rng, m, (-2.0, 2.0), half)

    eye = rng.uniform(-50, 50, 3)
    vp = perspective_vp(eye, -eye + rng.normal(scale=10.0, size=3), 60.0, 16 / 9, 0.1, 1000.0)
    return TreeScene(parent_slot, level_start.astype(np.uint64), local, index_of_slot, lower, upper, obj_slot, vp)


# ---------------------------------------------------------------------------
# batched multi-view
# ---------------------------------------------------------------------------
@dataclass
class BatchedScene:
    lower: np.ndarray
    upper: np.ndarray
    matrices: np.ndarray            # (n, 4, 4), object i uses matrix i
    group_first: np.ndarray         # (G,) u64
    group_count: np.ndarray         # (G,) u64
    vps: np.ndarray                 # (F, 4, 4) f32
    kinds: list = field(default_factory=list)  # per frustum: "cascade" | "cube" | "camera"
    pair_frustum: np.ndarray = None  # (P,) u32
    pair_group: np.ndarray = None    # (P,) u32
    pair_offset: np.ndarray = None   # (P,) u64 mask word offsets

    @property
    def n(self) -> int:
        return self.lower.shape[0]

    @property
    def mask_words(self) -> int:
        words = (self.group_count[self.pair_group] + 63) // 64
        return int((self.pair_offset + words).max()) if self.pair_group.size else 0

    @property
    def tests(self) -> int:
        return int(self.group_count[self.pair_group].sum())


def log_uniform_sizes(rng, n_groups, total, lo, hi) -> np.ndarray:
    raw = np.exp(rng.uniform(np.log(lo), np.log(hi), n_groups))
    sizes = np.maximum(1, np.floor(raw / raw.sum() * total)).astype(np.int64)
    sizes[np.argmax(sizes)] += total - sizes.sum()
    return sizes


def batched_scene(seed: int = 3, n: int = 2_097_152, n_groups: int = 64, n_cascades: int = 4,
                  n_cube: int = 6, n_cameras: int = 54, size_range=(1_000, 500_000),
                  groups_per_frustum=(1, 8), extent: float = 1000.0) -> BatchedScene:
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(rng, n_groups, n, size_range[0], size_range[1])
    first = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
    lower, upper = random_boxes(rng, n, (-extent, extent))
    mats = trs_matrices(rng, n)

    vps, kinds = [], []
    # Shadow cascades: one light direction, nested ortho boxes around a camera.
    light = _normalize(rng.normal(size=3))
    cam = rng.uniform(-extent / 20, extent / 20, 3)
    for c in range(n_cascades):
        half_w = 50.0 * (4.0 ** c)
        view = make_lookat(cam - light * 2 * extent, cam, _up_for(light))
        vps.append(view_projection(view, make_ortho(-half_w, half_w, -half_w, half_w, 0.1, 4 * extent)))
        kinds.append("cascade")
    # Cubemap: six 90-degree faces at one point.
    centre = rng.uniform(-extent / 4, extent / 4, 3)
    for d in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1])[:n_cube]:
        vps.append(perspective_vp(centre, np.array(d, np.float64), 90.0, 1.0, 0.1, 2 * extent))
        kinds.append("cube")
    for _ in range(n_cameras):
        eye = rng.uniform(-extent / 2, extent / 2, 3)
        vps.append(perspective_vp(eye, rng.normal(size=3), rng.uniform(30, 90), 16 / 9, 0.1,
                                  rng.uniform(100, 3000)))
        kinds.append("camera")
    vps = np.stack(vps).astype(np.float32)

    pf, pg = [], []
    for f in range(vps.shape[0]):
        k = int(rng.integers(groups_per_frustum[0], groups_per_frustum[1] + 1))
        for g in sorted(rng.choice(n_groups, size=min(k, n_groups), replace=False)):
            pf.append(f)
            pg.append(int(g))
    pf = np.array(pf, np.uint32)
    pg = np.array(pg, np.uint32)
    words = (sizes[pg] + 63) // 64
    off = np.concatenate([[0], np.cumsum(words)[:-1]]).astype(np.uint64)
    return BatchedScene(lower, upper, mats, first, sizes.astype(np.uint64), vps, kinds, pf, pg, off)


# ---------------------------------------------------------------------------
# serving: resident scenes and a request stream
# ---------------------------------------------------------------------------
@dataclass
class ServingScene:
    lower: np.ndarray      # (n, 3)
    upper: np.ndarray      # (n, 3)
    matrices: np.ndarray   # (n, 4, 4), object i uses matrix i
    perm: np.ndarray       # (n,) u32, a permutation: delta index sets are slices of it (distinct indices)
    pool: np.ndarray       # (ceil(n/10), 4, 4) replacement matrices: delta matrices are slices of it

    @property
    def n(self) -> int:
        return self.lower.shape[0]


@dataclass
class ServingRequest:
    scene: int
    vps: np.ndarray        # (F, 4, 4) f32
    delta_idx: np.ndarray  # (k,) u32 view into the scene's perm
    delta_mats: np.ndarray  # (k, 4, 4) f32 view into the scene's pool
    arrival_us: float      # offset from the start of the stream


def serving_scenes(seed: int = 4, n_scenes: int = 16, size_range=(10_000, 2_000_000),
                   extent: float = 1000.0) -> list:
    """Scenes of log-uniform size with the flat 1M-object distributions (box
    centres in [-extent, extent]^3, rotation x scale [0.5,2] x translation [-5,5]^3)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(size_range[0]), np.log(size_range[1]), n_scenes)).astype(np.int64)
    scenes = []
    for n in sizes:
        n = int(n)
        lower, upper = random_boxes(rng, n, (-extent, extent))
        mats = trs_matrices(rng, n)
        perm = rng.permutation(n).astype(np.uint32)
        pool = trs_matrices(rng, (n + 9) // 10)
        scenes.append(ServingScene(lower, upper, mats, perm, pool))
    return scenes


def serving_requests(scenes: list, seed: int, n_requests: int, rate_per_s: float, frusta=(1, 32),
                     delta_fraction=(0.01, 0.10), extent: float = 1000.0) -> list:
    """Poisson arrivals at `rate_per_s`; each request names a scene, carries 1-32
    random perspective frusta (fov 45, far 2000, eye in [-50,50]^3, random
    direction) and a matrix delta covering 1-10% of the scene's objects."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1e6 / rate_per_s, n_requests) if rate_per_s > 0 else np.zeros(n_requests)
    t = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n_requests):
        s = int(rng.integers(0, len(scenes)))
        sc = scenes[s]
        f = int(rng.integers(frusta[0], frusta[1] + 1))
        vps = np.stack([perspective_vp(rng.uniform(-50, 50, 3), rng.normal(size=3), 45.0, 16 / 9, 0.1, 2 * extent)
                        for _ in range(f)]).astype(np.float32)
        k = max(1, int(round(rng.uniform(*delta_fraction) * sc.n)))
        k = min(k, sc.pool.shape[0], sc.n)
        o1 = int(rng.integers(0, sc.n - k + 1))
        o2 = int(rng.integers(0, sc.pool.shape[0] - k + 1))
        out.append(ServingRequest(s, vps, sc.perm[o1:o1 + k], sc.pool[o2:o2 + k], float(t[i])))
    return out

// batch.hip -- many frusta over many object groups in one persistent launch,
// and the plain per-view form of the same work.
//
// Reference: cull(group, result, viewProjection) (dp/culling/Manager.h:139),
// one call per (view, group) in the reference (dp/culling/cpu/src/ManagerImpl.cpp:467-518).
//
// Work decomposition: a group is split into tiles of 64*K objects.  A work item
// is (tile, slice of the frusta paired with the tile's group).  A wave owns one
// item at a time: every lane builds the OBBs of its K objects once (one read of
// each object's matrix and box), then loops over the item's frusta.  Each
// frustum's matrix is wave-uniform, so it is read through the scalar cache into
// SGPRs and feeds the VALU multiplies directly; each (frustum, 64 objects)
// result is one ballot, stored as one 64-bit mask word by lane 0.  No two items
// write the same mask word, so no atomics touch the output.
//
// Schedule: items are sorted by cost (frusta x objects) on the host, and the
// persistent grid walks them round-robin (wave w takes w, w+W, ...), which
// bounds the load imbalance by the cost of one item.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch_types.hpp"
#include "cull_math.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace dpc {
namespace {

struct WorkItem {
  uint64_t first_object;
  uint32_t count;       // objects in the tile (<= 64*K)
  uint32_t tile_word;   // mask word of the tile's first object, relative to its group
  uint32_t pair_begin;  // slice of the pair table
  uint32_t pair_count;
  uint32_t scene;       // entry of the scene table
  uint32_t pad;
};
static_assert(sizeof(WorkItem) == 32, "WorkItem layout");


struct ViewTile {
  uint64_t first_object;
  uint64_t mask_word;
  uint32_t count;
  uint32_t scene;
  uint64_t pad;
};
static_assert(sizeof(ViewTile) == 32, "ViewTile layout");

// Device copy of a dpc_scene_desc (same layout).
struct SceneDev {
  const float4* box_a;
  const float2* box_b;
  const uint32_t* xform;
  const float4* planes;
  uint64_t ld;
  uint64_t n_objects;
};
static_assert(sizeof(SceneDev) == 48 && sizeof(SceneDev) == sizeof(dpc_scene_desc), "SceneDev layout");

__device__ __forceinline__ void load_vp(const float* __restrict__ vps, uint32_t frustum, Mat16& vp)
{
#pragma unroll
  for (int c = 0; c < 16; ++c) vp.m[c] = vps[16 * static_cast<size_t>(frustum) + c];
}

template <int K>
__device__ __forceinline__ void load_item(const SceneDev* __restrict__ scenes, const WorkItem& item, uint32_t lane,
                                          Slot (&slot)[K], RawObj (&raw)[K])
{
  const SceneDev sc = scenes[item.scene];
  const size_t end = item.first_object + item.count;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    slot[k] = make_slot(item.first_object + lane + kWave * k, end, sc.xform, sc.ld);
    raw[k] = load_raw(sc.box_a, sc.box_b, sc.planes, sc.ld, slot[k].i, slot[k].m);
  }
}

// The wave-uniform loop state is software-pipelined twice: the objects of the
// next work item are loaded (vector memory) while the current item is tested,
// and the next frustum matrix is loaded (scalar memory) while the current one is
// in use.  Both prefetches are unconditional (indices clamped to the last item
// or frustum), so no wait lands in front of the FP32 work.
// Tests a work item's objects (slot, o) against its frusta and stores one mask
// word per frustum and 64 objects.  The frusta take turns in two sets of scalar
// registers: while one frustum is tested, the next one's matrix is loaded into
// the other set (no register copies between frusta).  Every lane stores the
// wave's mask word (the same value to the same address): a store from lane 0
// alone would change the exec mask around it, scalar instructions in a loop
// whose scalar work is a bound.
template <int K>
__device__ __forceinline__ void test_frusta(const WorkItem& item, const Slot (&slot)[K], const Obb (&o)[K],
                                            const KPair* __restrict__ pairs, const float* __restrict__ vps,
                                            unsigned long long* __restrict__ masks)
{
  auto test = [&](const KPair& q, const Mat16& m) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool vis = slot_visible(slot[k], obb_visible(o[k], m.m));
      const unsigned long long bits = __ballot(vis);
      // (a work item holds at least one object: the first word is always stored)
      if (k == 0 || static_cast<uint32_t>(kWave * k) < item.count) masks[q.mask_word + item.tile_word + k] = bits;
    }
  };
  const uint32_t last = item.pair_count ? item.pair_count - 1 : 0;  // (no item is made without pairs)
  auto pair_at = [&](uint32_t p) { return pairs[item.pair_begin + (p < last ? p : last)]; };
  KPair pa = pair_at(0), pb;
  Mat16 va, vb;
  load_vp(vps, pa.frustum, va);
  for (uint32_t p = 0; p < item.pair_count; p += 2) {
    pb = pair_at(p + 1);
    load_vp(vps, pb.frustum, vb);
    test(pa, va);
    if (p + 1 >= item.pair_count) break;
    pa = pair_at(p + 2);
    load_vp(vps, pa.frustum, va);
    test(pb, vb);
  }
}

template <int K>
__global__ __launch_bounds__(256) void cull_batch_kernel(const WorkItem* __restrict__ items, uint32_t n_items,
                                                         const KPair* __restrict__ pairs,
                                                         const SceneDev* __restrict__ scenes,
                                                         const float* __restrict__ vps,
                                                         unsigned long long* __restrict__ masks)
{
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t waves_per_block = blockDim.x / kWave;
  const uint32_t n_waves = gridDim.x * waves_per_block;
  uint32_t it = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + threadIdx.x / kWave);
  if (it >= n_items) return;
  WorkItem item = items[it];
  Slot slot[K];
  RawObj raw[K];
  load_item<K>(scenes, item, lane, slot, raw);
  for (;;) {
    const uint32_t nit = it + n_waves;
    const WorkItem next_item = items[nit < n_items ? nit : it];
    Slot next_slot[K];
    RawObj next_raw[K];
    load_item<K>(scenes, next_item, lane, next_slot, next_raw);
    Obb o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = obb_of(raw[k]);
    test_frusta<K>(item, slot, o, pairs, vps, masks);
    if (nit >= n_items) break;
    it = nit;
    item = next_item;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      slot[k] = next_slot[k];
      raw[k] = next_raw[k];
    }
  }
}

// ---- serving form -----------------------------------------------------------
// One entry per request of a serving batch: the whole scene is tested against
// the request's frusta (pairs[pair_begin .. +pair_count)).  Work is split
// without a per-batch item list: a tile of entry a costs pair_count(a), every
// wave takes the tiles whose cost prefix falls in its equal share of the total
// cost, and walks them in order (the entry changes at most n_entries times).

template <int K>
__device__ __forceinline__ WorkItem serve_item(const ServeEntry& e, uint32_t tile)
{
  WorkItem w;
  w.first_object = static_cast<uint64_t>(tile) * (kWave * K);
  const uint64_t left = e.objects - w.first_object;
  w.count = static_cast<uint32_t>(left < static_cast<uint64_t>(kWave * K) ? left : static_cast<uint64_t>(kWave * K));
  w.tile_word = tile * K;
  w.pair_begin = e.pair_begin;
  w.pair_count = e.pair_count;
  w.scene = e.scene;
  w.pad = 0;
  return w;
}

template <int K>
__global__ __launch_bounds__(256) void cull_serve_kernel(const ServeEntry* __restrict__ entries, uint32_t n_entries,
                                                         uint64_t total_cost, const KPair* __restrict__ pairs,
                                                         const SceneDev* __restrict__ scenes,
                                                         const float* __restrict__ vps,
                                                         unsigned long long* __restrict__ masks)
{
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t waves_per_block = blockDim.x / kWave;
  const uint64_t W = static_cast<uint64_t>(gridDim.x) * waves_per_block;
  const uint64_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + threadIdx.x / kWave);
  const uint64_t c0 = total_cost * wave / W, c1 = total_cost * (wave + 1) / W;
  if (c0 >= c1) return;
  // first tile whose cost start is >= c0
  uint32_t a = 0;
  while (a + 1 < n_entries && entries[a + 1].cost_begin <= c0) ++a;
  ServeEntry e = entries[a];
  uint64_t t = e.pair_count ? (c0 - e.cost_begin + e.pair_count - 1) / e.pair_count : e.tiles;
  auto advance = [&](ServeEntry& ent, uint32_t& idx, uint64_t& tile) {
    while (tile >= ent.tiles && idx + 1 < n_entries) {
      ++idx;
      ent = entries[idx];
      tile = 0;
    }
  };
  advance(e, a, t);
  auto cost_start = [](const ServeEntry& ent, uint64_t tile) { return ent.cost_begin + tile * ent.pair_count; };
  if (t >= e.tiles || cost_start(e, t) >= c1) return;
  WorkItem item = serve_item<K>(e, static_cast<uint32_t>(t));
  Slot slot[K];
  RawObj raw[K];
  load_item<K>(scenes, item, lane, slot, raw);
  for (;;) {
    // the next tile of this wave (or the current one again, as a harmless prefetch target)
    ServeEntry ne = e;
    uint32_t na = a;
    uint64_t nt = t + 1;
    advance(ne, na, nt);
    const bool more = nt < ne.tiles && cost_start(ne, nt) < c1;
    const WorkItem next_item = more ? serve_item<K>(ne, static_cast<uint32_t>(nt)) : item;
    Slot next_slot[K];
    RawObj next_raw[K];
    load_item<K>(scenes, next_item, lane, next_slot, next_raw);
    Obb o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = obb_of(raw[k]);
    test_frusta<K>(item, slot, o, pairs, vps, masks);
    if (!more) break;
    e = ne;
    a = na;
    t = nt;
    item = next_item;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      slot[k] = next_slot[k];
      raw[k] = next_raw[k];
    }
  }
}

// Plain per-view launch: one wave per 64-object tile, the frustum by value.
__global__ __launch_bounds__(256) void cull_tiles_kernel(const ViewTile* __restrict__ tiles, uint32_t n_tiles,
                                                         const SceneDev* __restrict__ scenes, Mat16 vp,
                                                         unsigned long long* __restrict__ masks)
{
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
  if (t >= n_tiles) return;
  const ViewTile tile = tiles[t];
  const SceneDev sc = scenes[tile.scene];
  const Slot sl = make_slot(tile.first_object + lane, tile.first_object + tile.count, sc.xform, sc.ld);
  const bool vis = slot_visible(sl, obb_visible(obb_of(load_raw(sc.box_a, sc.box_b, sc.planes, sc.ld, sl.i, sl.m)), vp.m));
  const unsigned long long bits = __ballot(vis);
  if (lane == 0) masks[tile.mask_word] = bits;
}

}  // namespace
}  // namespace dpc

struct dpc_batch_plan {
  dpc_batch_config cfg;
  uint32_t n_frusta = 0;
  uint64_t mask_words = 0;
  uint64_t tests = 0;
  uint32_t n_items = 0;
  dpc::WorkItem* d_items = nullptr;
  dpc::KPair* d_pairs = nullptr;
  // groups read by at least one pair, for validating bound scenes
  std::vector<dpc_group_range> used_groups;
  uint32_t n_scenes_needed = 0;
  dpc::SceneDev* d_scenes = nullptr;
  uint32_t n_scenes_bound = 0;
  // per-view form
  std::vector<uint32_t> view_tile_begin;  // n_frusta + 1
  dpc::ViewTile* d_view_tiles = nullptr;
  int device = 0;
};

using namespace dpc;

namespace {

int validate_config(const dpc_batch_config& c)
{
  if (c.objects_per_lane != 1 && c.objects_per_lane != 2 && c.objects_per_lane != 4)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "batch config: objects_per_lane must be 1, 2 or 4");
  if (c.block_threads != 64 && c.block_threads != 128 && c.block_threads != 256)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "batch config: block_threads must be 64, 128 or 256");
  if (c.waves_per_cu == 0 || c.waves_per_cu > 64)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "batch config: waves_per_cu must be 1..64");
  return DPC_OK;
}

void free_plan(dpc_batch_plan* p)
{
  if (!p) return;
  if (p->d_items) (void)hipFree(p->d_items);
  if (p->d_pairs) (void)hipFree(p->d_pairs);
  if (p->d_view_tiles) (void)hipFree(p->d_view_tiles);
  if (p->d_scenes) (void)hipFree(p->d_scenes);
  delete p;
}

}  // namespace

extern "C" {

int dpc_batch_plan_create(const dpc_group_range* groups, size_t n_groups, const dpc_cull_pair* pairs, size_t n_pairs,
                          uint32_t n_frusta, const dpc_batch_config* cfg, dpc_batch_plan** out)
{
  if (!out) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_create: null output");
  *out = nullptr;
  if ((n_groups && !groups) || (n_pairs && !pairs))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_create: null groups/pairs");
  dpc_batch_config c{2, 256, 8, 0};
  if (cfg) c = *cfg;
  int rc = validate_config(c);
  if (rc) return rc;
  for (size_t k = 0; k < n_pairs; ++k) {
    if (pairs[k].frustum >= n_frusta || pairs[k].group >= n_groups)
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_create: pair %zu names frustum %u / group %u out of range",
                       k, pairs[k].frustum, pairs[k].group);
  }
  for (size_t g = 0; g < n_groups; ++g)
    if (groups[g].count > 0xFFFFFFFFull)
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_create: group %zu has more than 2^32 objects", g);

  auto* plan = new dpc_batch_plan();
  plan->cfg = c;
  plan->n_frusta = n_frusta;
  (void)hipGetDevice(&plan->device);

  // Pairs grouped by group, frusta ascending within a group.
  std::vector<uint32_t> order(n_pairs);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (pairs[a].group != pairs[b].group) return pairs[a].group < pairs[b].group;
    return pairs[a].frustum < pairs[b].frustum;
  });
  std::vector<KPair> kp(n_pairs);
  std::vector<uint64_t> group_pair_begin(n_groups + 1, 0);
  for (size_t k = 0; k < n_pairs; ++k) {
    const dpc_cull_pair& p = pairs[order[k]];
    kp[k] = KPair{p.frustum, 0u, p.mask_word_offset};
    group_pair_begin[p.group + 1]++;
  }
  for (size_t g = 0; g < n_groups; ++g) group_pair_begin[g + 1] += group_pair_begin[g];

  const uint64_t tile_objects = static_cast<uint64_t>(kWave) * c.objects_per_lane;
  std::vector<WorkItem> items;
  for (size_t g = 0; g < n_groups; ++g) {
    const uint64_t pb = group_pair_begin[g], pe = group_pair_begin[g + 1];
    if (pb == pe || groups[g].count == 0) continue;
    const uint64_t chunk = c.max_frusta_per_item ? c.max_frusta_per_item : (pe - pb);
    for (uint64_t t0 = 0; t0 < groups[g].count; t0 += tile_objects) {
      const uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(tile_objects, groups[g].count - t0));
      for (uint64_t q = pb; q < pe; q += chunk) {
        WorkItem w{};
        w.scene = groups[g].scene;
        w.first_object = groups[g].first_object + t0;
        w.count = cnt;
        w.tile_word = static_cast<uint32_t>(t0 / kWave);
        w.pair_begin = static_cast<uint32_t>(q);
        w.pair_count = static_cast<uint32_t>(std::min<uint64_t>(chunk, pe - q));
        items.push_back(w);
      }
    }
  }
  // Largest items first (cost = frusta x 64-object words).
  std::stable_sort(items.begin(), items.end(), [](const WorkItem& a, const WorkItem& b) {
    const uint64_t ca = static_cast<uint64_t>(a.pair_count) * ((a.count + 63) / 64);
    const uint64_t cb = static_cast<uint64_t>(b.pair_count) * ((b.count + 63) / 64);
    return ca > cb;
  });
  if (items.size() > 0xFFFFFFFFull) {
    free_plan(plan);
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_create: too many work items");
  }
  plan->n_items = static_cast<uint32_t>(items.size());

  // Per-view tiles: frustum-major, 64 objects per tile.
  std::vector<std::vector<ViewTile>> per_view(n_frusta);
  for (size_t k = 0; k < n_pairs; ++k) {
    const dpc_cull_pair& p = pairs[k];
    const dpc_group_range& g = groups[p.group];
    const uint64_t words = (g.count + 63) / 64;
    plan->mask_words = std::max(plan->mask_words, p.mask_word_offset + words);
    plan->tests += g.count;
    for (uint64_t w = 0; w < words; ++w) {
      ViewTile vt{};
      vt.first_object = g.first_object + 64 * w;
      vt.mask_word = p.mask_word_offset + w;
      vt.count = static_cast<uint32_t>(std::min<uint64_t>(64, g.count - 64 * w));
      vt.scene = g.scene;
      per_view[p.frustum].push_back(vt);
    }
  }
  std::vector<ViewTile> flat;
  plan->view_tile_begin.assign(n_frusta + 1, 0);
  for (uint32_t f = 0; f < n_frusta; ++f) {
    plan->view_tile_begin[f] = static_cast<uint32_t>(flat.size());
    flat.insert(flat.end(), per_view[f].begin(), per_view[f].end());
  }
  plan->view_tile_begin[n_frusta] = static_cast<uint32_t>(flat.size());
  {
    std::vector<bool> used(n_groups, false);
    for (size_t k = 0; k < n_pairs; ++k) used[pairs[k].group] = true;
    for (size_t g = 0; g < n_groups; ++g) {
      if (!used[g] || groups[g].count == 0) continue;
      plan->used_groups.push_back(groups[g]);
      plan->n_scenes_needed = std::max(plan->n_scenes_needed, groups[g].scene + 1);
    }
  }

  hipError_t e = hipSuccess;
  if (!items.empty()) {
    e = hipMalloc(&plan->d_items, items.size() * sizeof(WorkItem));
    if (e == hipSuccess)
      e = hipMemcpy(plan->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !kp.empty()) {
    e = hipMalloc(&plan->d_pairs, kp.size() * sizeof(KPair));
    if (e == hipSuccess) e = hipMemcpy(plan->d_pairs, kp.data(), kp.size() * sizeof(KPair), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !flat.empty()) {
    e = hipMalloc(&plan->d_view_tiles, flat.size() * sizeof(ViewTile));
    if (e == hipSuccess)
      e = hipMemcpy(plan->d_view_tiles, flat.data(), flat.size() * sizeof(ViewTile), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    free_plan(plan);
    return hip_fail(e, "dpc_batch_plan_create: upload");
  }
  *out = plan;
  return DPC_OK;
}

int dpc_batch_plan_destroy(dpc_batch_plan* plan)
{
  free_plan(plan);
  return DPC_OK;
}

int dpc_batch_plan_info(const dpc_batch_plan* plan, uint64_t* mask_words, uint64_t* tests, uint64_t* work_items)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_info: null plan");
  if (mask_words) *mask_words = plan->mask_words;
  if (tests) *tests = plan->tests;
  if (work_items) *work_items = plan->n_items;
  return DPC_OK;
}

}  // extern "C"

namespace {
int check_bound(const dpc_batch_plan* plan, size_t mask_words, const char* who)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: null plan", who);
  if (plan->n_scenes_bound < plan->n_scenes_needed)
    return set_error(DPC_ERR_STATE, "%s: the plan reads %u scenes but %u are bound (dpc_batch_plan_bind)", who,
                     plan->n_scenes_needed, plan->n_scenes_bound);
  if (mask_words < plan->mask_words)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: mask buffer holds %zu words, the pairs need %llu", who,
                     mask_words, static_cast<unsigned long long>(plan->mask_words));
  return DPC_OK;
}
}  // namespace

extern "C" {

int dpc_batch_plan_bind(dpc_batch_plan* plan, const dpc_scene_desc* scenes, size_t n_scenes)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_bind: null plan");
  if (n_scenes && !scenes) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_bind: null scene table");
  if (n_scenes < plan->n_scenes_needed)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_bind: the plan reads %u scenes, %zu given",
                     plan->n_scenes_needed, n_scenes);
  for (size_t k = 0; k < n_scenes; ++k) {
    const dpc_scene_desc& d = scenes[k];
    if (d.n_objects && (!d.box_a || !d.box_b || !d.planes || !aligned(d.box_a, 16) || !aligned(d.box_b, 8) ||
                        !aligned(d.planes, 16)))
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_bind: scene %zu has a null or misaligned array", k);
  }
  for (const dpc_group_range& g : plan->used_groups) {
    const dpc_scene_desc& d = scenes[g.scene];
    if (g.first_object + g.count > d.n_objects)
      return set_error(DPC_ERR_INVALID_ARGUMENT,
                       "dpc_batch_plan_bind: a group reaches object %llu of scene %u, which holds %llu",
                       static_cast<unsigned long long>(g.first_object + g.count), g.scene,
                       static_cast<unsigned long long>(d.n_objects));
    if (!d.xform && d.ld < g.first_object + g.count)
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_batch_plan_bind: scene %u has no indices and ld < objects",
                       g.scene);
  }
  std::vector<SceneDev> table(n_scenes);
  for (size_t k = 0; k < n_scenes; ++k) {
    table[k].box_a = reinterpret_cast<const float4*>(scenes[k].box_a);
    table[k].box_b = reinterpret_cast<const float2*>(scenes[k].box_b);
    table[k].xform = scenes[k].xform;
    table[k].planes = reinterpret_cast<const float4*>(scenes[k].planes);
    table[k].ld = scenes[k].ld;
    table[k].n_objects = scenes[k].n_objects;
  }
  if (n_scenes > plan->n_scenes_bound || !plan->d_scenes) {
    if (plan->d_scenes) (void)hipFree(plan->d_scenes);
    plan->d_scenes = nullptr;
    plan->n_scenes_bound = 0;
    DPC_HIP_TRY(hipMalloc(&plan->d_scenes, std::max<size_t>(n_scenes, 1) * sizeof(SceneDev)), "hipMalloc scenes");
  }
  if (n_scenes)
    DPC_HIP_TRY(hipMemcpy(plan->d_scenes, table.data(), n_scenes * sizeof(SceneDev), hipMemcpyHostToDevice),
                "upload scene table");
  plan->n_scenes_bound = static_cast<uint32_t>(n_scenes);
  return DPC_OK;
}

int dpc_cull_batch(const dpc_batch_plan* plan, const float* d_vps, uint64_t* d_masks, size_t mask_words,
                   void* stream)
{
  if (int rc = check_bound(plan, mask_words, "dpc_cull_batch")) return rc;
  if (plan->n_items == 0) return DPC_OK;
  if (!d_vps || !d_masks || !aligned(d_masks, 8) || !aligned(d_vps, 4))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_batch: null or misaligned frusta/masks");
  const dpc_batch_config& c = plan->cfg;
  const uint32_t waves_per_block = c.block_threads / kWave;
  uint64_t blocks = static_cast<uint64_t>(device_cu_count()) * c.waves_per_cu / waves_per_block;
  const uint64_t needed = (plan->n_items + waves_per_block - 1) / waves_per_block;
  if (blocks > needed) blocks = needed;
  if (blocks == 0) blocks = 1;
  const dim3 grid(static_cast<unsigned>(blocks)), block(c.block_threads);
  auto* m = reinterpret_cast<unsigned long long*>(d_masks);
  hipStream_t s = as_stream(stream);
  switch (c.objects_per_lane) {
    case 1:
      hipLaunchKernelGGL(cull_batch_kernel<1>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs,
                         plan->d_scenes, d_vps, m);
      break;
    case 2:
      hipLaunchKernelGGL(cull_batch_kernel<2>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs,
                         plan->d_scenes, d_vps, m);
      break;
    default:
      hipLaunchKernelGGL(cull_batch_kernel<4>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs,
                         plan->d_scenes, d_vps, m);
      break;
  }
  return check_launch("cull_batch_kernel");
}

}  // extern "C"

namespace dpc {
// Launches one serving batch (see ServeEntry).  d_entries/d_pairs/d_vps/d_masks
// live on the device; total_cost and n_entries come from the host copy.
// Copy-out of a serving batch's masks: blockIdx.y = the request, threads take
// consecutive 64-bit words (a wave writes 512 contiguous bytes per store).
// The destinations are pinned host memory, which device stores reach directly.
__global__ __launch_bounds__(256) void copy_out_kernel(const CopyOut* __restrict__ copies, const char* __restrict__ arena)
{
  const CopyOut c = copies[blockIdx.y];
  const auto* src = reinterpret_cast<const unsigned long long*>(arena + c.src_offset);
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; w < c.words;
       w += static_cast<uint64_t>(gridDim.x) * 256)
    c.dst[w] = src[w];
}

int launch_copy_out(const void* d_copies, uint32_t n, const char* d_arena, uint64_t max_words, void* stream_handle)
{
  if (n == 0 || max_words == 0) return DPC_OK;
  if (n > 65535) return set_error(DPC_ERR_INVALID_ARGUMENT, "launch_copy_out: more than 65535 ranges");
  // enough blocks per range for its words, at most 64 (then each thread loops)
  const uint64_t per = (max_words + 255) / 256;
  const unsigned bx = static_cast<unsigned>(per < 64 ? per : 64);
  hipLaunchKernelGGL(copy_out_kernel, dim3(bx, n), dim3(256), 0, as_stream(stream_handle),
                     static_cast<const CopyOut*>(d_copies), d_arena);
  return check_launch("copy_out_kernel");
}

int launch_serve(const void* d_entries, uint32_t n_entries, uint64_t total_cost, const void* d_pairs,
                 const void* d_scenes, const float* d_vps, uint64_t* d_masks, uint32_t objects_per_lane,
                 uint32_t waves_per_cu, void* stream_handle)
{
  hipStream_t stream = as_stream(stream_handle);
  if (n_entries == 0 || total_cost == 0) return DPC_OK;
  const uint32_t waves_per_block = 4;
  uint64_t waves = static_cast<uint64_t>(device_cu_count()) * (waves_per_cu ? waves_per_cu : 24);
  if (waves > total_cost) waves = total_cost;
  const uint64_t blocks = (waves + waves_per_block - 1) / waves_per_block;
  const auto* e = static_cast<const ServeEntry*>(d_entries);
  const auto* p = static_cast<const KPair*>(d_pairs);
  const auto* sc = static_cast<const SceneDev*>(d_scenes);
  auto* m = reinterpret_cast<unsigned long long*>(d_masks);
  switch (objects_per_lane) {
    case 2:
      hipLaunchKernelGGL(cull_serve_kernel<2>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, e, n_entries,
                         total_cost, p, sc, d_vps, m);
      break;
    default:
      hipLaunchKernelGGL(cull_serve_kernel<1>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, e, n_entries,
                         total_cost, p, sc, d_vps, m);
      break;
  }
  return check_launch("cull_serve_kernel");
}
}  // namespace dpc

extern "C" {

int dpc_cull_per_view(const dpc_batch_plan* plan, const float* h_vps, uint64_t* d_masks, size_t mask_words,
                      void* stream)
{
  if (int rc = check_bound(plan, mask_words, "dpc_cull_per_view")) return rc;
  if (!h_vps || !d_masks || !aligned(d_masks, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_per_view: null or misaligned frusta/masks");
  auto* m = reinterpret_cast<unsigned long long*>(d_masks);
  hipStream_t s = as_stream(stream);
  for (uint32_t f = 0; f < plan->n_frusta; ++f) {
    const uint32_t t0 = plan->view_tile_begin[f], t1 = plan->view_tile_begin[f + 1];
    if (t0 == t1) continue;
    Mat16 vp;
    std::memcpy(vp.m, h_vps + 16 * static_cast<size_t>(f), sizeof vp.m);
    const uint32_t n = t1 - t0;
    const unsigned blocks = (n + 3) / 4;
    hipLaunchKernelGGL(cull_tiles_kernel, dim3(blocks), dim3(256), 0, s, plan->d_view_tiles + t0, n, plan->d_scenes,
                       vp, m);
    int rc = check_launch("cull_tiles_kernel");
    if (rc) return rc;
  }
  return DPC_OK;
}

}  // extern "C"

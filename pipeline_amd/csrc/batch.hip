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
  uint32_t pad0, pad1;
};
static_assert(sizeof(WorkItem) == 32, "WorkItem layout");

struct KPair {
  uint32_t frustum;
  uint32_t pad;
  uint64_t mask_word;  // first mask word of the (frustum, group) result
};
static_assert(sizeof(KPair) == 16, "KPair layout");

struct ViewTile {
  uint64_t first_object;
  uint64_t mask_word;
  uint32_t count;
  uint32_t pad0;
  uint64_t pad1;
};
static_assert(sizeof(ViewTile) == 32, "ViewTile layout");

__device__ __forceinline__ Obb load_obb(const float4* __restrict__ box_a, const float2* __restrict__ box_b,
                                        const uint32_t* __restrict__ xform, const float4* __restrict__ planes,
                                        size_t ld, size_t i)
{
  const size_t m = xform ? static_cast<size_t>(xform[i]) : i;
  if (m >= ld) {
    // A matrix index outside the array is never dereferenced.  A NaN box makes
    // every clip coordinate NaN, so no cull flag is set and the object stays visible.
    const float q = __builtin_nanf("");
    Obb o;
    o.p = make_float4(q, q, q, q);
    o.ex = o.p;
    o.ey = o.p;
    o.ez = o.p;
    return o;
  }
  return make_obb(box_a[i], box_b[i], planes[m], planes[ld + m], planes[2 * ld + m], planes[3 * ld + m]);
}

template <int K>
__global__ __launch_bounds__(256) void cull_batch_kernel(const WorkItem* __restrict__ items, uint32_t n_items,
                                                         const KPair* __restrict__ pairs,
                                                         const float4* __restrict__ box_a,
                                                         const float2* __restrict__ box_b,
                                                         const uint32_t* __restrict__ xform,
                                                         const float4* __restrict__ planes, size_t ld,
                                                         const float* __restrict__ vps,
                                                         unsigned long long* __restrict__ masks)
{
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t waves_per_block = blockDim.x / kWave;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + threadIdx.x / kWave);
  const uint32_t n_waves = gridDim.x * waves_per_block;
  for (uint32_t it = wave; it < n_items; it += n_waves) {
    const WorkItem item = items[it];
    Obb o[K];
    bool valid[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t j = lane + kWave * k;
      valid[k] = j < item.count;
      if (valid[k]) o[k] = load_obb(box_a, box_b, xform, planes, ld, item.first_object + j);
    }
    for (uint32_t p = 0; p < item.pair_count; ++p) {
      const KPair pr = pairs[item.pair_begin + p];
      Mat16 vp;
#pragma unroll
      for (int c = 0; c < 16; ++c) vp.m[c] = vps[16 * static_cast<size_t>(pr.frustum) + c];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool vis = valid[k] && obb_visible(o[k], vp.m);
        const unsigned long long bits = __ballot(vis);
        if (lane == 0 && static_cast<uint32_t>(kWave * k) < item.count) masks[pr.mask_word + item.tile_word + k] = bits;
      }
    }
  }
}

// Plain per-view launch: one wave per 64-object tile, the frustum by value.
__global__ __launch_bounds__(256) void cull_tiles_kernel(const ViewTile* __restrict__ tiles, uint32_t n_tiles,
                                                         const float4* __restrict__ box_a,
                                                         const float2* __restrict__ box_b,
                                                         const uint32_t* __restrict__ xform,
                                                         const float4* __restrict__ planes, size_t ld, Mat16 vp,
                                                         unsigned long long* __restrict__ masks)
{
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
  if (t >= n_tiles) return;
  const ViewTile tile = tiles[t];
  bool vis = false;
  if (lane < tile.count) vis = obb_visible(load_obb(box_a, box_b, xform, planes, ld, tile.first_object + lane), vp.m);
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
  uint64_t object_end = 0;  // one past the last object any pair reads
  uint32_t n_items = 0;
  dpc::WorkItem* d_items = nullptr;
  dpc::KPair* d_pairs = nullptr;
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
    plan->object_end = std::max(plan->object_end, g.first_object + g.count);
    plan->tests += g.count;
    for (uint64_t w = 0; w < words; ++w) {
      ViewTile vt{};
      vt.first_object = g.first_object + 64 * w;
      vt.mask_word = p.mask_word_offset + w;
      vt.count = static_cast<uint32_t>(std::min<uint64_t>(64, g.count - 64 * w));
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
int check_extents(const dpc_batch_plan* plan, size_t n_objects, const uint32_t* d_xform, size_t ld, size_t mask_words,
                  const char* who)
{
  if (plan->object_end > n_objects)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: groups reach object %llu but the arrays hold %zu", who,
                     static_cast<unsigned long long>(plan->object_end), n_objects);
  if (!d_xform && ld < plan->object_end)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: objects without indices need ld >= %llu", who,
                     static_cast<unsigned long long>(plan->object_end));
  if (mask_words < plan->mask_words)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: mask buffer holds %zu words, the pairs need %llu", who,
                     mask_words, static_cast<unsigned long long>(plan->mask_words));
  return DPC_OK;
}
}  // namespace

extern "C" {

int dpc_cull_batch(const dpc_batch_plan* plan, const float* d_box_a, const float* d_box_b, const uint32_t* d_xform,
                   size_t n_objects, const float* d_planes, size_t ld, const float* d_vps, uint64_t* d_masks,
                   size_t mask_words, void* stream)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_batch: null plan");
  if (plan->n_items == 0) return DPC_OK;
  if (int rc = check_extents(plan, n_objects, d_xform, ld, mask_words, "dpc_cull_batch")) return rc;
  if (!d_box_a || !d_box_b || !d_planes || !d_vps || !d_masks)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_batch: null pointer argument");
  if (!aligned(d_box_a, 16) || !aligned(d_box_b, 8) || !aligned(d_planes, 16) || !aligned(d_masks, 8) ||
      !aligned(d_vps, 4))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_batch: misaligned pointer");
  const dpc_batch_config& c = plan->cfg;
  const uint32_t waves_per_block = c.block_threads / kWave;
  uint64_t blocks = static_cast<uint64_t>(device_cu_count()) * c.waves_per_cu / waves_per_block;
  const uint64_t needed = (plan->n_items + waves_per_block - 1) / waves_per_block;
  if (blocks > needed) blocks = needed;
  if (blocks == 0) blocks = 1;
  const dim3 grid(static_cast<unsigned>(blocks)), block(c.block_threads);
  const auto* a = reinterpret_cast<const float4*>(d_box_a);
  const auto* b = reinterpret_cast<const float2*>(d_box_b);
  const auto* p = reinterpret_cast<const float4*>(d_planes);
  auto* m = reinterpret_cast<unsigned long long*>(d_masks);
  hipStream_t s = as_stream(stream);
  switch (c.objects_per_lane) {
    case 1:
      hipLaunchKernelGGL(cull_batch_kernel<1>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs, a, b,
                         d_xform, p, ld, d_vps, m);
      break;
    case 2:
      hipLaunchKernelGGL(cull_batch_kernel<2>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs, a, b,
                         d_xform, p, ld, d_vps, m);
      break;
    default:
      hipLaunchKernelGGL(cull_batch_kernel<4>, grid, block, 0, s, plan->d_items, plan->n_items, plan->d_pairs, a, b,
                         d_xform, p, ld, d_vps, m);
      break;
  }
  return check_launch("cull_batch_kernel");
}

int dpc_cull_per_view(const dpc_batch_plan* plan, const float* d_box_a, const float* d_box_b, const uint32_t* d_xform,
                      size_t n_objects, const float* d_planes, size_t ld, const float* h_vps, uint64_t* d_masks,
                      size_t mask_words, void* stream)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_per_view: null plan");
  if (int rc = check_extents(plan, n_objects, d_xform, ld, mask_words, "dpc_cull_per_view")) return rc;
  if (!d_box_a || !d_box_b || !d_planes || !h_vps || !d_masks)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_per_view: null pointer argument");
  if (!aligned(d_box_a, 16) || !aligned(d_box_b, 8) || !aligned(d_planes, 16) || !aligned(d_masks, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_per_view: misaligned pointer");
  const auto* a = reinterpret_cast<const float4*>(d_box_a);
  const auto* b = reinterpret_cast<const float2*>(d_box_b);
  const auto* p = reinterpret_cast<const float4*>(d_planes);
  auto* m = reinterpret_cast<unsigned long long*>(d_masks);
  hipStream_t s = as_stream(stream);
  for (uint32_t f = 0; f < plan->n_frusta; ++f) {
    const uint32_t t0 = plan->view_tile_begin[f], t1 = plan->view_tile_begin[f + 1];
    if (t0 == t1) continue;
    Mat16 vp;
    std::memcpy(vp.m, h_vps + 16 * static_cast<size_t>(f), sizeof vp.m);
    const uint32_t n = t1 - t0;
    const unsigned blocks = (n + 3) / 4;
    hipLaunchKernelGGL(cull_tiles_kernel, dim3(blocks), dim3(256), 0, s, plan->d_view_tiles + t0, n, a, b, d_xform, p,
                       ld, vp, m);
    int rc = check_launch("cull_tiles_kernel");
    if (rc) return rc;
  }
  return DPC_OK;
}

}  // extern "C"

// tree.hip -- level-synchronous world-matrix propagation.
//
// Reference: dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166):
//   for each level, for each (parent, t) in the level's list:
//     if dirtyWorld[parent] || dirtyLocal[t]: world[t] = local[t] * world[parent]; dirtyWorld[t] = 1
//
// Device layout: local and world matrices in row planes, indexed by slot; the
// slots of one level are contiguous, in the reference's list order, so the
// loads of L[s] and the stores of W[s] are contiguous per wave and the parent
// rows W[parent[s]] are a short, monotone run of the previous level (children
// of one parent are adjacent).  Levels are dependent, so each level is one
// launch; runs of small levels share one single-workgroup launch that steps
// through them with a workgroup barrier in between.
#include <vector>

#include "cull_math.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace dpc {
namespace {

constexpr unsigned kTreeThreads = 256;
constexpr unsigned kSmallThreads = 1024;
constexpr uint64_t kSmallLevel = 8192;   // levels up to this size may share a single-workgroup launch
constexpr int kMaxLevelsPerSmall = 32;

struct LevelRun {
  uint64_t off[kMaxLevelsPerSmall + 1];
  int n;
};

__device__ __forceinline__ void propagate_slot(const float4* __restrict__ L, float4* W, size_t ld,
                                               const uint32_t* __restrict__ parent, uint64_t s,
                                               const uint8_t* __restrict__ dirty_local, uint8_t* dirty_world)
{
  const size_t p = parent[s];
  if (p >= ld) return;  // a parent outside the arrays is never read
  const bool dirty = (dirty_local == nullptr) || dirty_local[s] != 0 || dirty_world[p] != 0;
  if (!dirty) return;
  const float4 p0 = W[p], p1 = W[ld + p], p2 = W[2 * ld + p], p3 = W[3 * ld + p];
  const float4 l0 = L[s], l1 = L[ld + s], l2 = L[2 * ld + s], l3 = L[3 * ld + s];
  W[s] = rowmat(l0, p0, p1, p2, p3);
  W[ld + s] = rowmat(l1, p0, p1, p2, p3);
  W[2 * ld + s] = rowmat(l2, p0, p1, p2, p3);
  W[3 * ld + s] = rowmat(l3, p0, p1, p2, p3);
  dirty_world[s] = 1;
}

__global__ __launch_bounds__(kTreeThreads) void tree_level_kernel(const float4* __restrict__ L, float4* W, size_t ld,
                                                                  const uint32_t* __restrict__ parent, uint64_t s0,
                                                                  uint64_t s1, const uint8_t* __restrict__ dirty_local,
                                                                  uint8_t* dirty_world)
{
  const uint64_t s = s0 + static_cast<uint64_t>(blockIdx.x) * kTreeThreads + threadIdx.x;
  if (s < s1) propagate_slot(L, W, ld, parent, s, dirty_local, dirty_world);
}

// One workgroup walks several small levels; the barrier orders level l's stores
// before level l+1's loads (all waves of a workgroup share one CU and its L1).
__global__ __launch_bounds__(kSmallThreads) void tree_small_levels_kernel(const float4* __restrict__ L, float4* W,
                                                                          size_t ld,
                                                                          const uint32_t* __restrict__ parent,
                                                                          LevelRun run,
                                                                          const uint8_t* __restrict__ dirty_local,
                                                                          uint8_t* dirty_world)
{
  for (int l = 0; l < run.n; ++l) {
    for (uint64_t s = run.off[l] + threadIdx.x; s < run.off[l + 1]; s += kSmallThreads)
      propagate_slot(L, W, ld, parent, s, dirty_local, dirty_world);
    __syncthreads();
  }
}

__global__ void bytes_to_bits_kernel(const uint8_t* __restrict__ bytes, const uint32_t* __restrict__ index_of_slot,
                                     size_t n, uint32_t* __restrict__ bits, size_t n_bits)
{
  const size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n || bytes[s] == 0) return;
  const size_t i = index_of_slot ? static_cast<size_t>(index_of_slot[s]) : s;
  if (i >= n_bits) return;
  atomicOr(&bits[i >> 5], 1u << (i & 31));
}

}  // namespace
}  // namespace dpc

using namespace dpc;

extern "C" {

int dpc_tree_propagate(const float* d_local, float* d_world, size_t ld, const uint32_t* d_parent_slot,
                       const uint64_t* h_level_offsets, size_t n_levels, const uint8_t* d_dirty_local,
                       uint8_t* d_dirty_world, void* stream)
{
  if (n_levels == 0) return DPC_OK;
  if (!d_local || !d_world || !d_parent_slot || !h_level_offsets || !d_dirty_world)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate: null pointer argument");
  if (!aligned(d_local, 16) || !aligned(d_world, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate: misaligned matrix planes");
  for (size_t l = 0; l < n_levels; ++l)
    if (h_level_offsets[l] > h_level_offsets[l + 1])
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate: level offsets decrease at level %zu", l);
  if (h_level_offsets[n_levels] > ld)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate: levels end at slot %llu beyond ld=%zu",
                     static_cast<unsigned long long>(h_level_offsets[n_levels]), ld);
  const auto* L = reinterpret_cast<const float4*>(d_local);
  auto* W = reinterpret_cast<float4*>(d_world);
  hipStream_t s = as_stream(stream);
  size_t l = 0;
  while (l < n_levels) {
    const uint64_t size = h_level_offsets[l + 1] - h_level_offsets[l];
    if (size <= kSmallLevel) {
      LevelRun run;
      run.n = 0;
      run.off[0] = h_level_offsets[l];
      while (l < n_levels && run.n < kMaxLevelsPerSmall && h_level_offsets[l + 1] - h_level_offsets[l] <= kSmallLevel) {
        run.off[run.n + 1] = h_level_offsets[l + 1];
        ++run.n;
        ++l;
      }
      hipLaunchKernelGGL(tree_small_levels_kernel, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld, d_parent_slot, run,
                         d_dirty_local, d_dirty_world);
      int rc = check_launch("tree_small_levels_kernel");
      if (rc) return rc;
    } else {
      const unsigned blocks = static_cast<unsigned>((size + kTreeThreads - 1) / kTreeThreads);
      hipLaunchKernelGGL(tree_level_kernel, dim3(blocks), dim3(kTreeThreads), 0, s, L, W, ld, d_parent_slot,
                         h_level_offsets[l], h_level_offsets[l + 1], d_dirty_local, d_dirty_world);
      int rc = check_launch("tree_level_kernel");
      if (rc) return rc;
      ++l;
    }
  }
  return DPC_OK;
}

int dpc_bytes_to_bits_indexed(const uint8_t* d_bytes, const uint32_t* d_index_of_slot, size_t n_slots,
                              uint32_t* d_bits, size_t n_bits, void* stream)
{
  if (n_slots == 0 || n_bits == 0) return DPC_OK;
  if (!d_bytes || !d_bits) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_bytes_to_bits_indexed: null pointer");
  hipLaunchKernelGGL(bytes_to_bits_kernel, dim3(static_cast<unsigned>((n_slots + 255) / 256)), dim3(256), 0,
                     as_stream(stream), d_bytes, d_index_of_slot, n_slots, d_bits, n_bits);
  return check_launch("bytes_to_bits_kernel");
}

}  // extern "C"

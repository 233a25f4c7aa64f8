// tree.hip -- level-synchronous world-matrix propagation.
//
// Reference: dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166):
//   for each level, for each (parent, t) in the level's list:
//     if dirtyWorld[parent] || dirtyLocal[t]: world[t] = local[t] * world[parent]; dirtyWorld[t] = 1
//
// Device layout: local and world matrices in row planes, indexed by slot; the
// slots of one level are contiguous, in the reference's list order, so the
// loads of L[s] and the stores of W[s] are contiguous per wave.  Levels are
// dependent.  A large level is one launch.  A run of small levels (at most
// kLdsNodes slots each) shares one single-workgroup launch: the workgroup keeps
// the world matrices and dirty flags of the level it has just written in LDS,
// where the next level reads its parents, and loads the next level's local
// matrices and parent slots while it computes the current one, so each level
// costs a barrier and LDS traffic instead of dependent trips to memory.
#include <vector>

#include "cull_math.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace dpc {
namespace {

constexpr unsigned kTreeThreads = 256;
constexpr unsigned kSmallThreads = 1024;
constexpr unsigned kLdsNodes = 2048;  // levels up to this size may share a single-workgroup launch
constexpr int kPerThread = kLdsNodes / kSmallThreads;
constexpr int kMaxLevelsPerSmall = 32;
static_assert(kLdsNodes % kSmallThreads == 0, "slots per thread");

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

// What one thread reads of one slot before the parent is known.
struct SlotIn {
  float4 l0, l1, l2, l3;
  uint32_t p;
  uint32_t dl;
};

__device__ __forceinline__ void load_level(const float4* __restrict__ L, size_t ld, const uint32_t* __restrict__ parent,
                                           const uint8_t* __restrict__ dirty_local, uint64_t a, uint64_t b,
                                           SlotIn (&in)[kPerThread])
{
  if (a >= b) return;  // an empty level: nothing to read (a may equal ld)
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const uint64_t s = a + threadIdx.x + static_cast<uint64_t>(k) * kSmallThreads;
    const uint64_t c = s < b ? s : a;  // clamped: a slot of the level, read and then ignored
    in[k].p = parent[c];
    in[k].dl = dirty_local ? dirty_local[c] : 1u;
    in[k].l0 = L[c];
    in[k].l1 = L[ld + c];
    in[k].l2 = L[2 * ld + c];
    in[k].l3 = L[3 * ld + c];
  }
}

// One workgroup walks a run of small levels (each at most kLdsNodes slots; the
// host guarantees it).  LDS holds the previous level's world rows and dirty
// flags; a parent outside that level (the first level of the run, or a level
// split off after index reuse) is read from memory, where this workgroup or an
// earlier launch wrote it.  The barriers order one level's stores before the
// next level's loads (all waves of a workgroup share one CU).
__global__ __launch_bounds__(kSmallThreads) void tree_small_levels_kernel(const float4* __restrict__ L, float4* W,
                                                                          size_t ld,
                                                                          const uint32_t* __restrict__ parent,
                                                                          LevelRun run,
                                                                          const uint8_t* __restrict__ dirty_local,
                                                                          uint8_t* dirty_world)
{
  __shared__ float4 lds_w[4][kLdsNodes];
  __shared__ uint8_t lds_d[kLdsNodes];
  uint64_t prev0 = 0, prev1 = 0;  // the slots whose rows LDS holds
  SlotIn cur[kPerThread];
  load_level(L, ld, parent, dirty_local, run.off[0], run.off[1], cur);
  for (int l = 0; l < run.n; ++l) {
    const uint64_t a = run.off[l], b = run.off[l + 1];
    SlotIn nxt[kPerThread];
    if (l + 1 < run.n) load_level(L, ld, parent, dirty_local, run.off[l + 1], run.off[l + 2], nxt);
    float4 w[kPerThread][4];
    uint32_t dw[kPerThread];
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const uint64_t s = a + threadIdx.x + static_cast<uint64_t>(k) * kSmallThreads;
      dw[k] = 0;
      if (s >= b) continue;
      const uint64_t p = cur[k].p;
      if (p >= ld) {  // never read (as in propagate_slot); LDS gets the slot's current rows
        w[k][0] = W[s];
        w[k][1] = W[ld + s];
        w[k][2] = W[2 * ld + s];
        w[k][3] = W[3 * ld + s];
        dw[k] = dirty_world[s];
        continue;
      }
      const bool staged = p >= prev0 && p < prev1;
      const uint32_t dp = staged ? lds_d[p - prev0] : dirty_world[p];
      if (cur[k].dl | dp) {
        float4 p0, p1, p2, p3;
        if (staged) {
          p0 = lds_w[0][p - prev0];
          p1 = lds_w[1][p - prev0];
          p2 = lds_w[2][p - prev0];
          p3 = lds_w[3][p - prev0];
        } else {
          p0 = W[p];
          p1 = W[ld + p];
          p2 = W[2 * ld + p];
          p3 = W[3 * ld + p];
        }
        w[k][0] = rowmat(cur[k].l0, p0, p1, p2, p3);
        w[k][1] = rowmat(cur[k].l1, p0, p1, p2, p3);
        w[k][2] = rowmat(cur[k].l2, p0, p1, p2, p3);
        w[k][3] = rowmat(cur[k].l3, p0, p1, p2, p3);
        W[s] = w[k][0];
        W[ld + s] = w[k][1];
        W[2 * ld + s] = w[k][2];
        W[3 * ld + s] = w[k][3];
        dirty_world[s] = 1;
        dw[k] = 1;
      } else {  // clean: the world matrix is unchanged; its children may still need it
        w[k][0] = W[s];
        w[k][1] = W[ld + s];
        w[k][2] = W[2 * ld + s];
        w[k][3] = W[3 * ld + s];
        dw[k] = dirty_world[s];
      }
    }
    __syncthreads();  // every wave is done reading the previous level from LDS (and its stores are visible)
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) {
      const uint64_t s = a + threadIdx.x + static_cast<uint64_t>(k) * kSmallThreads;
      if (s >= b) continue;
      const uint64_t i = s - a;
      lds_w[0][i] = w[k][0];
      lds_w[1][i] = w[k][1];
      lds_w[2][i] = w[k][2];
      lds_w[3][i] = w[k][3];
      lds_d[i] = static_cast<uint8_t>(dw[k]);
    }
    prev0 = a;
    prev1 = b;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPerThread; ++k) cur[k] = nxt[k];
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
    if (size <= kLdsNodes) {
      LevelRun run;
      run.n = 0;
      run.off[0] = h_level_offsets[l];
      while (l < n_levels && run.n < kMaxLevelsPerSmall && h_level_offsets[l + 1] - h_level_offsets[l] <= kLdsNodes) {
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

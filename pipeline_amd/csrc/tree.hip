// tree.hip -- level-synchronous world-matrix propagation.
//
// Reference: dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166):
//   for each level, for each (parent, t) in the level's list:
//     if dirtyWorld[parent] || dirtyLocal[t]: world[t] = local[t] * world[parent]; dirtyWorld[t] = 1
//
// Device layout: local and world matrices in row planes, indexed by slot; the
// slots of one level are contiguous, so the loads of L[s] and the stores of
// W[s] are contiguous per wave.  Levels are
// dependent.  A large level is one launch.  A run of small levels (at most
// kLdsNodes slots each) shares one single-workgroup launch: the workgroup keeps
// the world matrices and dirty flags of the level it has just written in LDS,
// where the next level reads its parents, and loads the next level's local
// matrices and parent slots while it computes the current one, so each level
// costs a barrier and LDS traffic instead of dependent trips to memory.
#include <algorithm>
#include <memory>
#include <vector>

#include "cull_math.hpp"
#include "internal.hpp"

#pragma clang fp contract(off)

namespace dpc {
namespace {

constexpr unsigned kTreeThreads = 256;
constexpr unsigned kSmallThreads = 1024;
constexpr unsigned kLdsNodes = 1024;  // slots of one single-workgroup run (all of them staged in LDS)
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

template <bool HasDirty>
__device__ __forceinline__ void load_level(const float4* __restrict__ L, size_t ld, const uint32_t* __restrict__ parent,
                                           const uint8_t* __restrict__ dirty_local, uint64_t a, uint64_t b,
                                           SlotIn (&in)[kPerThread])
{
  // Unconditional and clamped: lanes past the level (and an empty level) read a
  // slot that exists (slot 0, the root, always does) and ignore it.  Loads
  // that every path issues let the compiler wait for exactly the loads a level
  // needs.
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const uint64_t s = a + threadIdx.x + static_cast<uint64_t>(k) * kSmallThreads;
    const uint64_t c = s < b ? s : (b > a ? a : 0);
    in[k].p = parent[c];
    in[k].dl = HasDirty ? dirty_local[c] : 1u;
    in[k].l0 = L[c];
    in[k].l1 = L[ld + c];
    in[k].l2 = L[2 * ld + c];
    in[k].l3 = L[3 * ld + c];
  }
}

// Orders one level's LDS stores before the next level's LDS loads.  Global
// memory needs no ordering inside a run: every parent the run reads from memory
// lies outside the slots the run writes before that read (it was written by an
// earlier launch, or is an unlisted slot or a slot of a later pass of the same
// level, whose old value the reference reads too), so the wait covers LDS only
// and the next level's prefetched loads stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A parent that is not staged in LDS, read from memory.  The empty asm takes
// the loaded values as register operands: it keeps the compiler from folding
// this load and the LDS read of the other branch into one generic-address load,
// and it places the wait for these loads inside this (rarely taken) branch
// rather than where the branches join.
__device__ __forceinline__ void load_parent(const float4* W, size_t ld, const uint8_t* dirty_world, uint64_t p,
                                            float4& p0, float4& p1, float4& p2, float4& p3, uint32_t& dp)
{
  p0 = W[p];
  p1 = W[ld + p];
  p2 = W[2 * ld + p];
  p3 = W[3 * ld + p];
  dp = dirty_world[p];
  asm volatile("" : "+v"(p0.x), "+v"(p0.y), "+v"(p0.z), "+v"(p0.w), "+v"(p1.x), "+v"(p1.y), "+v"(p1.z), "+v"(p1.w));
  asm volatile("" : "+v"(p2.x), "+v"(p2.y), "+v"(p2.z), "+v"(p2.w), "+v"(p3.x), "+v"(p3.y), "+v"(p3.z), "+v"(p3.w),
               "+v"(dp));
}

// Computes one level of a run from its prefetched inputs into LDS: the rows of
// every matrix the level recomputes and one flag per slot (1 = recomputed).  A
// parent outside the run, or one the run left unchanged, is read from memory,
// where it keeps its value for the whole launch.  Nothing is stored to memory
// here (tree_small_levels_kernel writes the run out at its end), so the only
// vector-memory operations in flight are loads, and waiting for the current
// level's inputs does not wait for the next level's prefetch.
__device__ __forceinline__ void run_level(const SlotIn (&in)[kPerThread], const float4* W, size_t ld,
                                          const uint8_t* dirty_world, uint64_t base, uint64_t a, uint64_t b,
                                          float4 (&lds_w)[4][kLdsNodes], uint8_t (&lds_d)[kLdsNodes])
{
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const uint64_t s = a + threadIdx.x + static_cast<uint64_t>(k) * kSmallThreads;
    if (s >= b) continue;
    const uint64_t p = in[k].p;
    const bool valid = p < ld;  // a parent outside the arrays is never read (as in propagate_slot)
    const bool staged = valid && p >= base && p < a;
    const uint64_t i = staged ? p - base : 0;
    const uint32_t dp_lds = staged ? lds_d[i] : 0u;
    bool dirty = false;
    float4 p0, p1, p2, p3;
    if (dp_lds) {  // the parent was recomputed in this run: its rows are in LDS
      p0 = lds_w[0][i];
      p1 = lds_w[1][i];
      p2 = lds_w[2][i];
      p3 = lds_w[3][i];
      dirty = true;
    } else if (valid && (!staged || in[k].dl)) {
      uint32_t dp = 0;
      load_parent(W, ld, dirty_world, p, p0, p1, p2, p3, dp);
      // a staged parent that the run left unchanged has dirty_world == 0
      dirty = (in[k].dl | (staged ? 0u : dp)) != 0;
    }
    const uint64_t j = s - base;
    if (dirty) {
      lds_w[0][j] = rowmat(in[k].l0, p0, p1, p2, p3);
      lds_w[1][j] = rowmat(in[k].l1, p0, p1, p2, p3);
      lds_w[2][j] = rowmat(in[k].l2, p0, p1, p2, p3);
      lds_w[3][j] = rowmat(in[k].l3, p0, p1, p2, p3);
    }
    lds_d[j] = dirty ? 1 : 0;
  }
}

// One workgroup walks a run of small levels whose slots total at most
// kLdsNodes (the host guarantees it).  Every matrix the run recomputes is kept
// in LDS, at its slot's offset in the run, where the later levels of the run
// read their parents; a parent outside the run comes from memory.  Levels go in
// pairs over two register sets, so that the loads of the next level are issued
// before the current one is computed and no register copy waits for them.  The
// recomputed rows and dirty flags go to memory once, after the last level.
template <bool HasDirty>
__global__ __launch_bounds__(kSmallThreads) void tree_small_levels_kernel(const float4* __restrict__ L, float4* W,
                                                                          size_t ld,
                                                                          const uint32_t* __restrict__ parent,
                                                                          LevelRun run,
                                                                          const uint8_t* __restrict__ dirty_local,
                                                                          uint8_t* dirty_world)
{
  __shared__ float4 lds_w[4][kLdsNodes];
  __shared__ uint8_t lds_d[kLdsNodes];
  const uint64_t base = run.off[0];
  const int last = run.n - 1;
  SlotIn A[kPerThread], B[kPerThread];
  load_level<HasDirty>(L, ld, parent, dirty_local, run.off[0], run.off[1], A);
  for (int l = 0; l < run.n; l += 2) {
    // the next level's inputs (past the last level: the last level again, unused)
    const int lb = l + 1 < run.n ? l + 1 : last;
    load_level<HasDirty>(L, ld, parent, dirty_local, run.off[lb], run.off[lb + 1], B);
    run_level(A, W, ld, dirty_world, base, run.off[l], run.off[l + 1], lds_w, lds_d);
    lds_barrier();
    if (l + 1 >= run.n) break;
    const int la = l + 2 < run.n ? l + 2 : last;
    load_level<HasDirty>(L, ld, parent, dirty_local, run.off[la], run.off[la + 1], A);
    run_level(B, W, ld, dirty_world, base, run.off[l + 1], run.off[l + 2], lds_w, lds_d);
    lds_barrier();
  }
  const uint64_t total = run.off[run.n] - base;
  for (uint64_t j = threadIdx.x; j < total; j += kSmallThreads) {
    if (!lds_d[j]) continue;
    const uint64_t s = base + j;
    W[s] = lds_w[0][j];
    W[ld + s] = lds_w[1][j];
    W[2 * ld + s] = lds_w[2][j];
    W[3 * ld + s] = lds_w[3][j];
    dirty_world[s] = 1;
  }
}

// ---- tile runs -------------------------------------------------------------
// A run of consecutive passes in which every entry's parent lies in the
// previous pass and each pass is ordered by parent slot.  Then the children of
// a contiguous slot range form a contiguous range of the next pass, and so on
// down the run: a tile is a contiguous range of the run's first pass together
// with those descendant ranges, at most kTileSlots slots per pass.  One
// workgroup walks one tile pass by pass, keeping the previous pass's
// recomputed rows and flags in LDS, where the next pass reads its parents.
// Every parent read from memory is unchanged during the launch (a parent of
// the run's first pass was written by an earlier launch; a parent the tile
// did not recompute keeps its value), so the passes are ordered by LDS-only
// barriers.
constexpr unsigned kTileSlots = 1024;  // the widest tile (one slot per thread)
constexpr unsigned kDefaultTileSlots = 512;  // two workgroups per CU; measured faster than 1024 on the 1M-node tree
constexpr int kMaxTileLevels = 16;

struct TileLevel {
  uint32_t first;  // first slot of the tile in this pass
  uint32_t count;  // slots of the tile in this pass (<= kTileSlots)
};

template <bool HasDirty>
__device__ __forceinline__ void load_tile_slot(const float4* __restrict__ L, size_t ld,
                                               const uint32_t* __restrict__ parent,
                                               const uint8_t* __restrict__ dirty_local, TileLevel t, SlotIn& in)
{
  // unconditional and clamped (slot 0 exists), as load_level
  const uint64_t c = threadIdx.x < t.count ? static_cast<uint64_t>(t.first) + threadIdx.x : 0;
  in.p = parent[c];
  in.dl = HasDirty ? dirty_local[c] : 1u;
  in.l0 = L[c];
  in.l1 = L[ld + c];
  in.l2 = L[2 * ld + c];
  in.l3 = L[3 * ld + c];
}

template <bool HasDirty, unsigned Slots>
__device__ __forceinline__ void tile_level(const SlotIn& in, float4* W, size_t ld, uint8_t* dirty_world, TileLevel t,
                                           TileLevel prev, bool first_pass, float4 (&lds_in)[4][Slots],
                                           uint8_t (&flag_in)[Slots], float4 (&lds_out)[4][Slots],
                                           uint8_t (&flag_out)[Slots])
{
  if (threadIdx.x >= t.count) return;
  const uint64_t s = static_cast<uint64_t>(t.first) + threadIdx.x;
  const uint64_t p = in.p;
  const bool valid = p < ld;  // a parent outside the arrays is never read (as in propagate_slot)
  // a parent of a later pass lies in the tile's previous range (by construction)
  const bool staged = valid && !first_pass && p >= prev.first && p < static_cast<uint64_t>(prev.first) + prev.count;
  const uint32_t i = staged ? static_cast<uint32_t>(p - prev.first) : 0u;
  const uint32_t dp_lds = staged ? flag_in[i] : 0u;
  bool dirty = false;
  float4 p0, p1, p2, p3;
  if (dp_lds) {
    p0 = lds_in[0][i];
    p1 = lds_in[1][i];
    p2 = lds_in[2][i];
    p3 = lds_in[3][i];
    dirty = true;
  } else if (valid && (!staged || in.dl)) {
    uint32_t dp = 0;
    load_parent(W, ld, dirty_world, p, p0, p1, p2, p3, dp);
    dirty = (in.dl | (staged ? 0u : dp)) != 0;
  }
  if (dirty) {
    const float4 w0 = rowmat(in.l0, p0, p1, p2, p3);
    const float4 w1 = rowmat(in.l1, p0, p1, p2, p3);
    const float4 w2 = rowmat(in.l2, p0, p1, p2, p3);
    const float4 w3 = rowmat(in.l3, p0, p1, p2, p3);
    W[s] = w0;
    W[ld + s] = w1;
    W[2 * ld + s] = w2;
    W[3 * ld + s] = w3;
    dirty_world[s] = 1;
    lds_out[0][threadIdx.x] = w0;
    lds_out[1][threadIdx.x] = w1;
    lds_out[2][threadIdx.x] = w2;
    lds_out[3][threadIdx.x] = w3;
  }
  flag_out[threadIdx.x] = dirty ? 1 : 0;
}

template <bool HasDirty, unsigned Slots>
__global__ __launch_bounds__(Slots) void tree_tile_kernel(const float4* __restrict__ L, float4* W, size_t ld,
                                                          const uint32_t* __restrict__ parent,
                                                          const TileLevel* __restrict__ tiles, int n_levels,
                                                          const uint8_t* __restrict__ dirty_local,
                                                          uint8_t* dirty_world)
{
  __shared__ float4 rows[2][4][Slots];
  __shared__ uint8_t flags[2][Slots];
  // the tile's ranges, read once: later reads come from LDS, so waiting for
  // them never waits for the vector-memory loads in flight
  __shared__ TileLevel t[kMaxTileLevels];
  if (threadIdx.x < static_cast<unsigned>(n_levels)) t[threadIdx.x] = tiles[static_cast<size_t>(blockIdx.x) * n_levels + threadIdx.x];
  __syncthreads();
  const int last = n_levels - 1;
  SlotIn A, B;
  load_tile_slot<HasDirty>(L, ld, parent, dirty_local, t[0], A);
  for (int l = 0; l < n_levels; l += 2) {
    load_tile_slot<HasDirty>(L, ld, parent, dirty_local, t[l + 1 < n_levels ? l + 1 : last], B);
    tile_level<HasDirty, Slots>(A, W, ld, dirty_world, t[l], t[l > 0 ? l - 1 : 0], l == 0, rows[1], flags[1],
                                rows[0], flags[0]);
    lds_barrier();
    if (l + 1 >= n_levels) break;
    load_tile_slot<HasDirty>(L, ld, parent, dirty_local, t[l + 2 < n_levels ? l + 2 : last], A);
    tile_level<HasDirty, Slots>(B, W, ld, dirty_world, t[l + 1], t[l], false, rows[0], flags[0], rows[1], flags[1]);
    lds_barrier();
  }
}

// dp::math::invertTranspose (dp/math/Matmnt.h:1171-1201) of W[s]: 2x2 minors of
// rows 0-1 (s*) and rows 2-3 (c*), the determinant, one correctly rounded
// division, then each cofactor sum times 1/det, evaluated left to right as the
// reference writes it.  Output row i, column j = the reference's mOut[i][j].
__device__ __forceinline__ bool invert_transpose(const float4 r0, const float4 r1, const float4 r2, const float4 r3,
                                                 float4& o0, float4& o1, float4& o2, float4& o3)
{
  const float m00 = r0.x, m01 = r0.y, m02 = r0.z, m03 = r0.w;
  const float m10 = r1.x, m11 = r1.y, m12 = r1.z, m13 = r1.w;
  const float m20 = r2.x, m21 = r2.y, m22 = r2.z, m23 = r2.w;
  const float m30 = r3.x, m31 = r3.y, m32 = r3.z, m33 = r3.w;
  const float s0 = m00 * m11 - m01 * m10, c5 = m22 * m33 - m23 * m32;
  const float s1 = m00 * m12 - m02 * m10, c4 = m21 * m33 - m23 * m31;
  const float s2 = m00 * m13 - m03 * m10, c3 = m21 * m32 - m22 * m31;
  const float s3 = m01 * m12 - m02 * m11, c2 = m20 * m33 - m23 * m30;
  const float s4 = m01 * m13 - m03 * m11, c1 = m20 * m32 - m22 * m30;
  const float s5 = m02 * m13 - m03 * m12, c0 = m20 * m31 - m21 * m30;
  const float det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
  if (!(det != 0.0f)) return false;
  const float inv = 1.0f / det;
  o0.x = ((m11 * c5 - m12 * c4) + m13 * c3) * inv;
  o1.x = ((-m01 * c5 + m02 * c4) - m03 * c3) * inv;
  o2.x = ((m31 * s5 - m32 * s4) + m33 * s3) * inv;
  o3.x = ((-m21 * s5 + m22 * s4) - m23 * s3) * inv;
  o0.y = ((-m10 * c5 + m12 * c2) - m13 * c1) * inv;
  o1.y = ((m00 * c5 - m02 * c2) + m03 * c1) * inv;
  o2.y = ((-m30 * s5 + m32 * s2) - m33 * s1) * inv;
  o3.y = ((m20 * s5 - m22 * s2) + m23 * s1) * inv;
  o0.z = ((m10 * c4 - m11 * c2) + m13 * c0) * inv;
  o1.z = ((-m00 * c4 + m01 * c2) - m03 * c0) * inv;
  o2.z = ((m30 * s4 - m31 * s2) + m33 * s0) * inv;
  o3.z = ((-m20 * s4 + m21 * s2) - m23 * s0) * inv;
  o0.w = ((-m10 * c3 + m11 * c1) - m12 * c0) * inv;
  o1.w = ((m00 * c3 - m01 * c1) + m02 * c0) * inv;
  o2.w = ((-m30 * s3 + m31 * s1) - m32 * s0) * inv;
  o3.w = ((m20 * s3 - m21 * s1) + m22 * s0) * inv;
  return true;
}

// One thread per slot: reads W[s] (4 coalesced row loads), writes the inverse
// transpose in the same row-plane layout.  64 B in + 64 B out per matrix.
__global__ __launch_bounds__(kTreeThreads) void inverse_transpose_kernel(const float4* __restrict__ W, size_t ld,
                                                                         const uint8_t* __restrict__ dirty, size_t n,
                                                                         float4* __restrict__ out, size_t out_ld,
                                                                         uint8_t* __restrict__ ok)
{
  const size_t s = static_cast<size_t>(blockIdx.x) * kTreeThreads + threadIdx.x;
  if (s >= n) return;
  if (dirty && dirty[s] == 0) return;
  float4 o0, o1, o2, o3;
  const bool good = invert_transpose(W[s], W[ld + s], W[2 * ld + s], W[3 * ld + s], o0, o1, o2, o3);
  if (good) {
    out[s] = o0;
    out[out_ld + s] = o1;
    out[2 * out_ld + s] = o2;
    out[3 * out_ld + s] = o3;
  }
  if (ok) ok[s] = good ? 1 : 0;
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

// The launch schedule of one tree layout (dpc_tree_plan_create).
struct dpc_tree_plan {
  enum Kind { kSmall, kLevel, kTiles };
  struct Step {
    Kind kind;
    LevelRun run;          // kSmall
    uint64_t s0 = 0, s1 = 0;  // kLevel
    int n_levels = 0;      // kTiles
    uint32_t n_tiles = 0;
    TileLevel* d_tiles = nullptr;
  };
  size_t ld = 0;
  uint32_t tile_slots = kTileSlots;
  uint32_t* d_parent = nullptr;
  std::vector<Step> steps;
  uint64_t tiles = 0;
  ~dpc_tree_plan()
  {
    for (auto& st : steps)
      if (st.d_tiles) (void)hipFree(st.d_tiles);
    if (d_parent) (void)hipFree(d_parent);
  }
};

namespace {

// Host side of the plan: which consecutive passes form tile runs, and their tiles.
struct PlanBuilder {
  const uint32_t* par;
  const uint64_t* off;
  size_t n;  // passes
  uint32_t width = kTileSlots;  // slots of a tile per pass
  std::vector<uint8_t> link;  // link[j]: every parent of pass j is in pass j-1 and pass j is ordered by parent

  PlanBuilder(const uint32_t* parent, const uint64_t* offsets, size_t n_levels) : par(parent), off(offsets), n(n_levels)
  {
    link.assign(n, 0);
    for (size_t j = 1; j < n; ++j) {
      bool ok = off[j + 1] > off[j];
      for (uint64_t s = off[j]; ok && s < off[j + 1]; ++s) {
        ok = par[s] >= off[j - 1] && par[s] < off[j];
        if (ok && s > off[j]) ok = par[s] >= par[s - 1];
      }
      link[j] = ok;
    }
  }

  // first[v - off[j-1]] for v in [off[j-1], off[j]]: the first slot of pass j
  // whose parent is >= v (pass j ordered by parent).
  std::vector<uint32_t> first_child(size_t j) const
  {
    const uint64_t a = off[j - 1], b = off[j];
    std::vector<uint32_t> f(b - a + 1);
    uint64_t c = off[j];
    for (uint64_t v = a; v <= b; ++v) {
      while (c < off[j + 1] && par[c] < v) ++c;
      f[v - a] = static_cast<uint32_t>(c);
    }
    return f;
  }

  // Descendant boundaries of the nodes of pass l: cb[k][x - off[l]] is the first
  // slot of pass l + k below node x (x = off[l + 1] gives the end of the pass).
  // Extends while every link holds and no single node has more than `width`
  // descendants in a pass; returns the number of passes covered.
  size_t descend(size_t l, std::vector<std::vector<uint32_t>>& cb) const
  {
    cb.clear();
    std::vector<uint32_t> cur(off[l + 1] - off[l] + 1);
    for (size_t x = 0; x < cur.size(); ++x) cur[x] = static_cast<uint32_t>(off[l] + x);
    cb.push_back(cur);
    for (size_t j = l + 1; j < n && cb.size() < static_cast<size_t>(kMaxTileLevels); ++j) {
      if (!link[j]) break;
      const std::vector<uint32_t> f = first_child(j);
      std::vector<uint32_t> next(cur.size());
      uint32_t widest = 0;
      for (size_t x = 0; x < cur.size(); ++x) {
        next[x] = f[cur[x] - off[j - 1]];
        if (x) widest = std::max(widest, next[x] - next[x - 1]);
      }
      if (widest > width) break;
      cb.push_back(next);
      cur.swap(next);
    }
    return cb.size();
  }

  // Tiles of the run of passes [l, l + cb.size()): greedy contiguous ranges of
  // pass l, each at most `width` slots in every pass.
  std::vector<TileLevel> tiles(size_t l, const std::vector<std::vector<uint32_t>>& cb) const
  {
    const size_t levels = cb.size();
    const size_t nodes = off[l + 1] - off[l];
    std::vector<TileLevel> out;
    size_t x0 = 0;
    while (x0 < nodes) {
      size_t x1 = x0 + 1;
      while (x1 < nodes) {
        bool fits = true;
        for (size_t k = 0; k < levels && fits; ++k) fits = cb[k][x1 + 1] - cb[k][x0] <= width;
        if (!fits) break;
        ++x1;
      }
      for (size_t k = 0; k < levels; ++k) out.push_back(TileLevel{cb[k][x0], cb[k][x1] - cb[k][x0]});
      x0 = x1;
    }
    return out;
  }
};

}  // namespace

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
      // a run: consecutive small levels while their slots total at most kLdsNodes
      LevelRun run;
      run.n = 0;
      run.off[0] = h_level_offsets[l];
      while (l < n_levels && run.n < kMaxLevelsPerSmall &&
             h_level_offsets[l + 1] - run.off[0] <= kLdsNodes) {
        run.off[run.n + 1] = h_level_offsets[l + 1];
        ++run.n;
        ++l;
      }
      if (d_dirty_local)
        hipLaunchKernelGGL(tree_small_levels_kernel<true>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld, d_parent_slot,
                           run, d_dirty_local, d_dirty_world);
      else
        hipLaunchKernelGGL(tree_small_levels_kernel<false>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld,
                           d_parent_slot, run, d_dirty_local, d_dirty_world);
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

int dpc_inverse_transpose(const float* d_planes, size_t ld, const uint8_t* d_dirty, size_t n, float* d_out_planes,
                          size_t out_ld, uint8_t* d_ok, void* stream)
{
  if (n == 0) return DPC_OK;
  if (!d_planes || !d_out_planes) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_inverse_transpose: null pointer");
  if (n > ld || n > out_ld)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_inverse_transpose: n=%zu exceeds ld=%zu or out_ld=%zu", n, ld, out_ld);
  if (!aligned(d_planes, 16) || !aligned(d_out_planes, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_inverse_transpose: misaligned matrix planes");
  hipLaunchKernelGGL(inverse_transpose_kernel, dim3(static_cast<unsigned>((n + kTreeThreads - 1) / kTreeThreads)),
                     dim3(kTreeThreads), 0, as_stream(stream), reinterpret_cast<const float4*>(d_planes), ld, d_dirty, n,
                     reinterpret_cast<float4*>(d_out_planes), out_ld, d_ok);
  return check_launch("inverse_transpose_kernel");
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

int dpc_tree_plan_create(const uint32_t* h_parent_slot, size_t ld, const uint64_t* h_level_offsets, size_t n_levels,
                         uint32_t tile_slots, dpc_tree_plan** out)
{
  if (!out) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: null plan pointer");
  *out = nullptr;
  if (!h_parent_slot || !h_level_offsets || n_levels == 0 || ld == 0 || ld > 0xFFFFFFFFull)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: bad arguments");
  for (size_t l = 0; l < n_levels; ++l)
    if (h_level_offsets[l] > h_level_offsets[l + 1])
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: level offsets decrease at level %zu", l);
  if (h_level_offsets[n_levels] > ld)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: levels end beyond ld");
  if (tile_slots != 0 && tile_slots != 512 && tile_slots != 1024)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: tile_slots must be 0, 512 or 1024");
  auto plan = std::make_unique<dpc_tree_plan>();
  plan->ld = ld;
  plan->tile_slots = tile_slots ? tile_slots : kDefaultTileSlots;
  DPC_HIP_TRY(hipMalloc(&plan->d_parent, ld * sizeof(uint32_t)), "dpc_tree_plan_create: parents");
  DPC_HIP_TRY(hipMemcpy(plan->d_parent, h_parent_slot, ld * sizeof(uint32_t), hipMemcpyHostToDevice),
              "dpc_tree_plan_create: parents");
  PlanBuilder pb(h_parent_slot, h_level_offsets, n_levels);
  pb.width = plan->tile_slots;
  const uint64_t* off = h_level_offsets;
  std::vector<std::vector<uint32_t>> cb, cb_next;
  size_t l = 0;
  while (l < n_levels) {
    const uint64_t size = off[l + 1] - off[l];
    if (size == 0) {
      ++l;
      continue;
    }
    if (size <= kLdsNodes) {  // consecutive small passes: one single-workgroup run
      dpc_tree_plan::Step st{};
      st.kind = dpc_tree_plan::kSmall;
      st.run.n = 0;
      st.run.off[0] = off[l];
      while (l < n_levels && st.run.n < kMaxLevelsPerSmall && off[l + 1] - st.run.off[0] <= kLdsNodes) {
        st.run.off[st.run.n + 1] = off[l + 1];
        ++st.run.n;
        ++l;
      }
      plan->steps.push_back(st);
      continue;
    }
    // a tile run, unless starting at the next pass reaches deeper
    const size_t reach = pb.descend(l, cb);
    bool tile = reach >= 2;
    if (tile && l + reach < n_levels && l + 1 < n_levels && off[l + 2] - off[l + 1] > kLdsNodes) {
      const size_t reach_next = pb.descend(l + 1, cb_next);
      if (l + 1 + reach_next > l + reach) tile = false;
    }
    if (tile) {
      pb.descend(l, cb);  // cb may have been reused above
      const std::vector<TileLevel> t = pb.tiles(l, cb);
      dpc_tree_plan::Step st{};
      st.kind = dpc_tree_plan::kTiles;
      st.n_levels = static_cast<int>(cb.size());
      st.n_tiles = static_cast<uint32_t>(t.size() / cb.size());
      DPC_HIP_TRY(hipMalloc(&st.d_tiles, t.size() * sizeof(TileLevel)), "dpc_tree_plan_create: tiles");
      plan->steps.push_back(st);
      DPC_HIP_TRY(hipMemcpy(st.d_tiles, t.data(), t.size() * sizeof(TileLevel), hipMemcpyHostToDevice),
                  "dpc_tree_plan_create: tiles");
      plan->tiles += st.n_tiles;
      l += cb.size();
      continue;
    }
    dpc_tree_plan::Step st{};
    st.kind = dpc_tree_plan::kLevel;
    st.s0 = off[l];
    st.s1 = off[l + 1];
    plan->steps.push_back(st);
    ++l;
  }
  *out = plan.release();
  return DPC_OK;
}

int dpc_tree_plan_destroy(dpc_tree_plan* plan)
{
  delete plan;
  return DPC_OK;
}

int dpc_tree_plan_info(const dpc_tree_plan* plan, uint32_t* launches, uint32_t* tile_runs, uint64_t* tiles)
{
  if (!plan) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_info: null plan");
  uint32_t runs = 0;
  for (const auto& st : plan->steps) runs += st.kind == dpc_tree_plan::kTiles ? 1 : 0;
  if (launches) *launches = static_cast<uint32_t>(plan->steps.size());
  if (tile_runs) *tile_runs = runs;
  if (tiles) *tiles = plan->tiles;
  return DPC_OK;
}

int dpc_tree_propagate_plan(const dpc_tree_plan* plan, const float* d_local, float* d_world,
                            const uint8_t* d_dirty_local, uint8_t* d_dirty_world, void* stream)
{
  if (!plan || !d_local || !d_world || !d_dirty_world)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate_plan: null argument");
  if (!aligned(d_local, 16) || !aligned(d_world, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_propagate_plan: misaligned matrix planes");
  const auto* L = reinterpret_cast<const float4*>(d_local);
  auto* W = reinterpret_cast<float4*>(d_world);
  const size_t ld = plan->ld;
  hipStream_t s = as_stream(stream);
  for (const auto& st : plan->steps) {
    switch (st.kind) {
      case dpc_tree_plan::kSmall:
        if (d_dirty_local)
          hipLaunchKernelGGL(tree_small_levels_kernel<true>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld,
                             plan->d_parent, st.run, d_dirty_local, d_dirty_world);
        else
          hipLaunchKernelGGL(tree_small_levels_kernel<false>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld,
                             plan->d_parent, st.run, d_dirty_local, d_dirty_world);
        break;
      case dpc_tree_plan::kLevel:
        hipLaunchKernelGGL(tree_level_kernel, dim3(static_cast<unsigned>((st.s1 - st.s0 + kTreeThreads - 1) / kTreeThreads)),
                           dim3(kTreeThreads), 0, s, L, W, ld, plan->d_parent, st.s0, st.s1, d_dirty_local,
                           d_dirty_world);
        break;
      case dpc_tree_plan::kTiles:
        if (plan->tile_slots == 512) {
          if (d_dirty_local)
            hipLaunchKernelGGL((tree_tile_kernel<true, 512>), dim3(st.n_tiles), dim3(512), 0, s, L, W, ld,
                               plan->d_parent, st.d_tiles, st.n_levels, d_dirty_local, d_dirty_world);
          else
            hipLaunchKernelGGL((tree_tile_kernel<false, 512>), dim3(st.n_tiles), dim3(512), 0, s, L, W, ld,
                               plan->d_parent, st.d_tiles, st.n_levels, d_dirty_local, d_dirty_world);
        } else {
          if (d_dirty_local)
            hipLaunchKernelGGL((tree_tile_kernel<true, 1024>), dim3(st.n_tiles), dim3(1024), 0, s, L, W, ld,
                               plan->d_parent, st.d_tiles, st.n_levels, d_dirty_local, d_dirty_world);
          else
            hipLaunchKernelGGL((tree_tile_kernel<false, 1024>), dim3(st.n_tiles), dim3(1024), 0, s, L, W, ld,
                               plan->d_parent, st.d_tiles, st.n_levels, d_dirty_local, d_dirty_world);
        }
        break;
    }
    if (int rc = check_launch("dpc_tree_propagate_plan")) return rc;
  }
  return DPC_OK;
}

}  // extern "C"

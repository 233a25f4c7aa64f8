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
#include "tree_plan.hpp"

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
// One launch per run of linked passes (tree_plan.hpp describes the tiles and
// their steps).  A workgroup takes one tile: step after step it loads one slot
// per thread (the step's pieces may span several passes), then computes the
// step's pieces in order; a kept pass's rows and flags go to the tile's LDS
// ring, where the next pass reads its parents, and an LDS barrier closes each
// kept pass.  The loads of the next step are issued before the current step is
// computed, into the other of two register sets.  Slots the tile owns are
// written to memory once per step, after its pieces.
using tileplan::Seg;
constexpr uint32_t kMaxSegs = 64;
constexpr uint32_t kNoSeg = 0xFFFFFFFFu;

struct StepIn {
  float4 l0, l1, l2, l3;
  uint32_t p;
  uint32_t dl;
  uint32_t seg;   // this thread's piece (kNoSeg: idle in this step)
  uint32_t slot;
  uint32_t row;   // the matrix row this thread computes in a kQuad piece (0 otherwise)
};

// A piece's fields as wave-uniform values (scalar registers): the first 16
// bytes of a Seg in one LDS load, and the parent range in another.
struct SegHead {
  uint32_t first, count, pos, ring, flags, own_end;
};

__device__ __forceinline__ SegHead decode_head(const uint4 h)
{
  SegHead r;
  r.first = __builtin_amdgcn_readfirstlane(h.x);
  const uint32_t cp = __builtin_amdgcn_readfirstlane(h.y);  // count | pos << 16
  const uint32_t rf = __builtin_amdgcn_readfirstlane(h.z);  // ring | flags << 16
  r.count = cp & 0xFFFFu;
  r.pos = cp >> 16;
  r.ring = rf & 0xFFFFu;
  r.flags = rf >> 16;
  r.own_end = __builtin_amdgcn_readfirstlane(h.w);
  return r;
}

__device__ __forceinline__ SegHead seg_head(const Seg* sg, uint32_t j)
{
  return decode_head(*reinterpret_cast<const uint4*>(&sg[j]));
}

// Finds this thread's piece among the pieces of the step that starts at piece
// g (four pieces per round of LDS loads) and issues its loads (unconditional
// and clamped, as load_level).  Returns the first piece of the next step (n
// when this is the last step).
template <bool HasDirty>
__device__ __forceinline__ uint32_t load_step(const float4* __restrict__ L, size_t ld,
                                              const uint32_t* __restrict__ parent,
                                              const uint8_t* __restrict__ dirty_local, const Seg* sg, uint32_t n,
                                              uint32_t g, StepIn& in)
{
  uint32_t mine = kNoSeg, slot = 0, row = 0;
  uint32_t next = g;
  bool ended = false;
  for (uint32_t j = g; !ended; j += 4) {
    SegHead h[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) h[q] = seg_head(sg, min(j + q, n - 1));
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      if (ended) continue;
      const uint32_t d = threadIdx.x - h[q].pos;
      const uint32_t sh = (h[q].flags & tileplan::kQuad) ? 2u : 0u;  // threads per slot: 1 << sh
      if (d < (h[q].count << sh)) {
        mine = j + q;
        slot = h[q].first + (d >> sh);
        row = d & ((1u << sh) - 1u);
      }
      next = j + q + 1;
      ended = (h[q].flags & tileplan::kStepEnd) || next >= n;
    }
  }
  next = min(next, n);
  in.seg = mine;
  in.slot = slot;
  in.row = row;
  const uint64_t c = mine != kNoSeg ? slot : 0;
  in.p = parent[c];
  in.dl = HasDirty ? dirty_local[c] : 1u;
  in.l0 = L[c];
  in.l1 = L[ld + c];
  in.l2 = L[2 * ld + c];
  in.l3 = L[3 * ld + c];
  return next;
}

// Buffer stores with an out-of-range offset are dropped by the hardware: a
// lane that must not write gets kNoStore, so every step issues the same store
// instructions on every path and the compiler can count the loads still in
// flight exactly (a store behind a branch makes it wait for all of them).
constexpr uint32_t kNoStore = 0x80000000u;
constexpr uint64_t kMaxBufferSlots = 1ull << 27;  // row planes addressed with 32-bit byte offsets below kNoStore
constexpr int kRawBufferDword3 = 0x00020000;      // gfx9 raw buffer: 32-bit data format, no swizzle

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void store_row(float4 v, __amdgpu_buffer_rsrc_t r, uint32_t offset)
{
  u32x4 u;
  u.x = __float_as_uint(v.x);
  u.y = __float_as_uint(v.y);
  u.z = __float_as_uint(v.z);
  u.w = __float_as_uint(v.w);
  __builtin_amdgcn_raw_buffer_store_b128(u, r, offset, 0, 0);
}

// Row in.row of the local matrix, selected with bit masks (exact for every
// bit pattern, and kept in registers: a select the compiler turns into an
// indexed access would put the four rows in scratch memory).
__device__ __forceinline__ float4 pick_row(const StepIn& in)
{
  const uint32_t m0 = in.row == 0 ? ~0u : 0u, m1 = in.row == 1 ? ~0u : 0u;
  const uint32_t m2 = in.row == 2 ? ~0u : 0u, m3 = in.row == 3 ? ~0u : 0u;
  auto pick = [&](float a, float b, float c, float d) {
    return __uint_as_float((__float_as_uint(a) & m0) | (__float_as_uint(b) & m1) | (__float_as_uint(c) & m2) |
                           (__float_as_uint(d) & m3));
  };
  return make_float4(pick(in.l0.x, in.l1.x, in.l2.x, in.l3.x), pick(in.l0.y, in.l1.y, in.l2.y, in.l3.y),
                     pick(in.l0.z, in.l1.z, in.l2.z, in.l3.z), pick(in.l0.w, in.l1.w, in.l2.w, in.l3.w));
}

// Computes the pieces [a, b) of one step.  Each thread has at most one slot in
// the step: its rows go to the ring (kept passes) inside the loop, and to
// memory (when the tile owns the slot and it changed) after the loop.
template <bool HasDirty, unsigned MaxRing>
__device__ __forceinline__ void run_step(const StepIn& in, const Seg* sg, uint32_t a, uint32_t b, uint32_t ring_len,
                                         const float4* W, size_t ld, const uint8_t* dirty_world,
                                         const __amdgpu_buffer_rsrc_t (&rw)[4], __amdgpu_buffer_rsrc_t rdw,
                                         float4 (&rows)[4][MaxRing], uint8_t (&flag)[MaxRing])
{
  bool store = false, quad = false;
  float4 w0 = make_float4(0.f, 0.f, 0.f, 0.f), w1 = w0, w2 = w0, w3 = w0;
  // a < b <= n for every step of a valid plan; b <= n always (load_step)
  const uint32_t a0 = a < b ? a : 0u;
  SegHead h = seg_head(sg, a0);
  uint2 pr = *reinterpret_cast<const uint2*>(&sg[a0].pfirst);  // the parent range: pfirst, pcount | pring << 16
  for (uint32_t j = a; j < b; ++j) {
    // the next piece's descriptor (both halves), read while this one is computed
    const uint32_t jn = j + 1 < b ? j + 1 : j;
    const uint4 hn = *reinterpret_cast<const uint4*>(&sg[jn]);
    const uint2 prn = *reinterpret_cast<const uint2*>(&sg[jn].pfirst);
    if (in.seg == j) {
      const uint32_t pfirst = __builtin_amdgcn_readfirstlane(pr.x);
      const uint32_t pp = __builtin_amdgcn_readfirstlane(pr.y);
      const uint32_t pcount = pp & 0xFFFFu, pring = pp >> 16;
      const uint64_t p = in.p;
      const bool valid = p < ld;  // a parent outside the arrays is never read (as in propagate_slot)
      const uint32_t d = static_cast<uint32_t>(p) - pfirst;
      const bool staged = valid && !(h.flags & tileplan::kFromMemory) && d < pcount;
      uint32_t i = pring + d;
      i = i >= ring_len ? i - ring_len : i;
      i = staged && i < MaxRing ? i : 0u;
      // the flag and the rows in one round of LDS loads (the rows are used
      // only when the flag says the parent was recomputed)
      const uint32_t f = flag[i];
      float4 p0 = rows[0][i], p1 = rows[1][i], p2 = rows[2][i], p3 = rows[3][i];
      const uint32_t dp_lds = staged ? f : 0u;
      bool dirty = false;
      if (dp_lds) {  // the parent was recomputed by this tile: its rows are in the ring
        dirty = true;
      } else if (valid && (!staged || in.dl)) {
        const uint32_t pu = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(p));
        if (!HasDirty && __builtin_amdgcn_ballot_w64(p != pu) == 0) {
          // every slot of the wave has the same parent outside the run (the
          // root, for a run from the top): its rows come through the scalar
          // cache, which leaves the vector loads of later steps in flight.  The
          // launch never writes that slot, and without dirty flags its
          // dirty_world is not needed.
          typedef const __attribute__((address_space(4))) float* ConstFloats;
          const ConstFloats wc = (ConstFloats)(reinterpret_cast<uintptr_t>(W));
          const size_t q = 4 * static_cast<size_t>(pu), r = 4 * ld;
          p0 = make_float4(wc[q], wc[q + 1], wc[q + 2], wc[q + 3]);
          p1 = make_float4(wc[r + q], wc[r + q + 1], wc[r + q + 2], wc[r + q + 3]);
          p2 = make_float4(wc[2 * r + q], wc[2 * r + q + 1], wc[2 * r + q + 2], wc[2 * r + q + 3]);
          p3 = make_float4(wc[3 * r + q], wc[3 * r + q + 1], wc[3 * r + q + 2], wc[3 * r + q + 3]);
          dirty = true;
        } else {
          // a kept parent that the tile left unchanged has dirty_world == 0
          uint32_t dp = 0;
          load_parent(W, ld, dirty_world, p, p0, p1, p2, p3, dp);
          dirty = (in.dl | (staged ? 0u : dp)) != 0;
        }
      }
      quad = (h.flags & tileplan::kQuad) != 0;
      if (dirty) {
        if (quad) {  // this thread's row only: a quarter of the arithmetic latency
          w0 = rowmat(pick_row(in), p0, p1, p2, p3);
          w1 = w0;
          w2 = w0;
          w3 = w0;
        } else {
          w0 = rowmat(in.l0, p0, p1, p2, p3);
          w1 = rowmat(in.l1, p0, p1, p2, p3);
          w2 = rowmat(in.l2, p0, p1, p2, p3);
          w3 = rowmat(in.l3, p0, p1, p2, p3);
        }
      }
      if (h.flags & tileplan::kStaged) {
        uint32_t o = h.ring + (in.slot - h.first);
        o = o >= ring_len ? o - ring_len : o;
        o = o < MaxRing ? o : 0u;  // the plan keeps positions below its ring length (at most MaxRing)
        if (dirty) {
          if (!quad || in.row == 0) rows[0][o] = w0;
          if (!quad || in.row == 1) rows[1][o] = w1;
          if (!quad || in.row == 2) rows[2][o] = w2;
          if (!quad || in.row == 3) rows[3][o] = w3;
        }
        flag[o] = dirty ? 1 : 0;
      }
      store = dirty && in.slot < h.own_end;
    }
    // passes written and read within one wave need no barrier: a wave's LDS
    // operations execute in order
    if (h.flags & tileplan::kBarrier) lds_barrier();
    h = decode_head(hn);
    pr = prn;
  }
  // one row plane per thread of a kQuad slot, all four otherwise
  const uint32_t off = store ? in.slot * 16u : kNoStore;
  store_row(w0, rw[0], !quad || in.row == 0 ? off : kNoStore);
  store_row(w1, rw[1], !quad || in.row == 1 ? off : kNoStore);
  store_row(w2, rw[2], !quad || in.row == 2 ? off : kNoStore);
  store_row(w3, rw[3], !quad || in.row == 3 ? off : kNoStore);
  __builtin_amdgcn_raw_buffer_store_b8(static_cast<unsigned char>(1), rdw, store && in.row == 0 ? in.slot : kNoStore,
                                       0, 0);
  // the step's loaded registers are consumed on every path: reusing them for a
  // later step's loads then waits for nothing
  asm volatile("" ::"v"(in.l0.x), "v"(in.l0.y), "v"(in.l0.z), "v"(in.l0.w), "v"(in.l1.x), "v"(in.l1.y),
               "v"(in.l1.z), "v"(in.l1.w));
  asm volatile("" ::"v"(in.l2.x), "v"(in.l2.y), "v"(in.l2.z), "v"(in.l2.w), "v"(in.l3.x), "v"(in.l3.y),
               "v"(in.l3.z), "v"(in.l3.w), "v"(in.p), "v"(in.dl));
}

// LDS ring positions per tile: four workgroups of 256 threads per CU (39 KB
// each), two of 512 (77 KB) or one of 1024 (133 KB).  The plan may use fewer.
constexpr unsigned tile_max_ring(unsigned threads)
{
  return threads == 256 ? 576 : threads == 512 ? 1152 : 2048;
}

template <bool HasDirty, unsigned Threads>
__global__ __launch_bounds__(Threads) void tree_tile_kernel(const float4* __restrict__ L, float4* W, size_t ld,
                                                            const uint32_t* __restrict__ parent,
                                                            const Seg* __restrict__ segs, uint32_t ring_len,
                                                            const uint8_t* __restrict__ dirty_local,
                                                            uint8_t* dirty_world)
{
  constexpr unsigned MaxRing = tile_max_ring(Threads);
  __shared__ float4 rows[4][MaxRing];
  __shared__ uint8_t flag[MaxRing];
  __shared__ Seg sg[kMaxSegs];
  // the tile's pieces: kMaxSegs per tile, the first one carrying the counts
  if (threadIdx.x < kMaxSegs) sg[threadIdx.x] = segs[static_cast<size_t>(blockIdx.x) * kMaxSegs + threadIdx.x];
  __syncthreads();
  const uint32_t n = min(__builtin_amdgcn_readfirstlane(sg[0].tile_pieces), kMaxSegs);
  const uint32_t steps = min(__builtin_amdgcn_readfirstlane(sg[0].tile_steps), n);
  if (n == 0) return;  // (the plan gives every tile at least one piece)
  const uint32_t rec = static_cast<uint32_t>(ld * 16);  // ld <= kMaxBufferSlots (checked by the plan)
  const __amdgpu_buffer_rsrc_t rw[4] = {__builtin_amdgcn_make_buffer_rsrc(W, 0, rec, kRawBufferDword3),
                                        __builtin_amdgcn_make_buffer_rsrc(W + ld, 0, rec, kRawBufferDword3),
                                        __builtin_amdgcn_make_buffer_rsrc(W + 2 * ld, 0, rec, kRawBufferDword3),
                                        __builtin_amdgcn_make_buffer_rsrc(W + 3 * ld, 0, rec, kRawBufferDword3)};
  const __amdgpu_buffer_rsrc_t rdw =
      __builtin_amdgcn_make_buffer_rsrc(dirty_world, 0, static_cast<uint32_t>(ld), kRawBufferDword3);
  // Steps rotate over three register sets, so each step's loads are issued two
  // steps before it is computed: the long first step (the tile's chain of small
  // passes) covers the loads of the next two.  The trip count is fixed before
  // the loop, so the loop has one exit and the compiler sees every load of a
  // set consumed before the set is loaded again.  A load past the last step
  // repeats an earlier step's (unused).
  StepIn A, B, C;
  uint32_t pa = 0;                                                                   // first piece of A's step
  uint32_t pb = load_step<HasDirty>(L, ld, parent, dirty_local, sg, n, pa, A);           // ... of B's
  uint32_t pc = load_step<HasDirty>(L, ld, parent, dirty_local, sg, n, pb < n ? pb : pa, B);  // ... of C's
  uint32_t k = 0;
  for (; k + 2 < steps; k += 3) {
    const uint32_t pd = load_step<HasDirty>(L, ld, parent, dirty_local, sg, n, pc, C);
    run_step<HasDirty, MaxRing>(A, sg, pa, pb, ring_len, W, ld, dirty_world, rw, rdw, rows, flag);
    const uint32_t pe = load_step<HasDirty>(L, ld, parent, dirty_local, sg, n, pd < n ? pd : pc, A);
    run_step<HasDirty, MaxRing>(B, sg, pb, pc, ring_len, W, ld, dirty_world, rw, rdw, rows, flag);
    const uint32_t pf = load_step<HasDirty>(L, ld, parent, dirty_local, sg, n, pe < n ? pe : pc, B);
    run_step<HasDirty, MaxRing>(C, sg, pc, pd, ring_len, W, ld, dirty_world, rw, rdw, rows, flag);
    pa = pd;
    pb = pe;
    pc = pf;
  }
  if (k < steps) run_step<HasDirty, MaxRing>(A, sg, pa, pb, ring_len, W, ld, dirty_world, rw, rdw, rows, flag);
  if (k + 1 < steps) run_step<HasDirty, MaxRing>(B, sg, pb, pc, ring_len, W, ld, dirty_world, rw, rdw, rows, flag);
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
  struct Step {
    tileplan::Step::Kind kind;
    LevelRun run{};            // kSmall
    uint64_t s0 = 0, s1 = 0;   // kLevel
    uint32_t n_tiles = 0;      // kTiles
    uint32_t ring = 0;
    Seg* d_segs = nullptr;
    };
  size_t ld = 0;
  uint32_t threads = 0;  // of a tile workgroup
  uint32_t* d_parent = nullptr;
  std::vector<Step> steps;
  uint64_t tiles = 0;
  ~dpc_tree_plan()
  {
    for (auto& st : steps) {
      if (st.d_segs) (void)hipFree(st.d_segs);
    }
    if (d_parent) (void)hipFree(d_parent);
  }
};

namespace {

constexpr uint32_t kDefaultTileThreads = 1024;  // measured fastest on the 1M-node tree (27.5 us vs 29.3 at 512)
constexpr size_t kMaxRunPasses = 32;

template <bool HasDirty>
void launch_tiles(const dpc_tree_plan& plan, const dpc_tree_plan::Step& st, const float4* L, float4* W,
                  const uint8_t* dirty_local, uint8_t* dirty_world, hipStream_t s)
{
  const dim3 grid(st.n_tiles);
  switch (plan.threads) {
    case 256:
      hipLaunchKernelGGL((tree_tile_kernel<HasDirty, 256>), grid, dim3(256), 0, s, L, W, plan.ld, plan.d_parent,
                         st.d_segs, st.ring, dirty_local, dirty_world);
      break;
    case 1024:
      hipLaunchKernelGGL((tree_tile_kernel<HasDirty, 1024>), grid, dim3(1024), 0, s, L, W, plan.ld, plan.d_parent,
                         st.d_segs, st.ring, dirty_local, dirty_world);
      break;
    default:
      hipLaunchKernelGGL((tree_tile_kernel<HasDirty, 512>), grid, dim3(512), 0, s, L, W, plan.ld, plan.d_parent,
                         st.d_segs, st.ring, dirty_local, dirty_world);
      break;
  }
}

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
  if (!h_parent_slot || !h_level_offsets || ld == 0 || ld > 0xFFFFFFFFull)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: bad arguments");
  for (size_t l = 0; l < n_levels; ++l)
    if (h_level_offsets[l] > h_level_offsets[l + 1])
      return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: level offsets decrease at level %zu", l);
  if (h_level_offsets[n_levels] > ld)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: levels end beyond ld");
  if (tile_slots != 0 && tile_slots != 256 && tile_slots != 512 && tile_slots != 1024)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_tree_plan_create: tile_slots must be 0, 256, 512 or 1024");
  if (n_levels == 0) {
    // a tree that holds only its root: propagation has nothing to compute
    // (Tree::compute visits no level and notifies an empty change set)
    auto plan = std::make_unique<dpc_tree_plan>();
    plan->ld = ld;
    plan->threads = tile_slots ? tile_slots : kDefaultTileThreads;
    *out = plan.release();
    return DPC_OK;
  }
  auto plan = std::make_unique<dpc_tree_plan>();
  plan->ld = ld;
  plan->threads = tile_slots ? tile_slots : kDefaultTileThreads;
  const uint32_t max_ring = tile_max_ring(plan->threads);
  // workgroups the device holds at once: CUs x tile workgroups per CU
  int dev = 0, cus = 0, per_cu = 0;
  DPC_HIP_TRY(hipGetDevice(&dev), "dpc_tree_plan_create: device");
  DPC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "dpc_tree_plan_create: device");
  switch (plan->threads) {
    case 256:
      DPC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tree_tile_kernel<true, 256>, 256, 0),
                  "dpc_tree_plan_create: occupancy");
      break;
    case 1024:
      DPC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tree_tile_kernel<true, 1024>, 1024, 0),
                  "dpc_tree_plan_create: occupancy");
      break;
    default:
      DPC_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tree_tile_kernel<true, 512>, 512, 0),
                  "dpc_tree_plan_create: occupancy");
      break;
  }
  const uint64_t capacity = static_cast<uint64_t>(std::max(cus, 1)) * static_cast<uint64_t>(std::max(per_cu, 1));
  DPC_HIP_TRY(hipMalloc(&plan->d_parent, ld * sizeof(uint32_t)), "dpc_tree_plan_create: parents");
  DPC_HIP_TRY(hipMemcpy(plan->d_parent, h_parent_slot, ld * sizeof(uint32_t), hipMemcpyHostToDevice),
              "dpc_tree_plan_create: parents");
  // a parent slot at or beyond ld is never read; the planner follows parents of
  // linked passes only, which lie inside the levels
  // tile runs address the row planes with 32-bit buffer offsets: none beyond kMaxBufferSlots slots
  const size_t run_passes = ld <= kMaxBufferSlots ? kMaxRunPasses : 1;
  const std::vector<tileplan::Step> sched =
      tileplan::schedule(h_parent_slot, h_level_offsets, n_levels, plan->threads, max_ring, capacity, kLdsNodes,
                         kMaxLevelsPerSmall, run_passes);
  for (const auto& ts : sched) {
    dpc_tree_plan::Step st{};
    st.kind = ts.kind;
    if (ts.kind == tileplan::Step::kSmall) {
      st.run.n = static_cast<int>(ts.r - ts.l);
      for (size_t k = ts.l; k <= ts.r; ++k) st.run.off[k - ts.l] = h_level_offsets[k];
    } else if (ts.kind == tileplan::Step::kLevel) {
      st.s0 = h_level_offsets[ts.l];
      st.s1 = h_level_offsets[ts.r];
    } else {
      const std::string why =
          ts.run.ring == 0 || ts.run.ring > max_ring
              ? std::string("ring larger than the kernel's LDS")
              : tileplan::check_run(h_parent_slot, h_level_offsets, tileplan::limits_for(plan->threads, ts.run.ring),
                                    ts.run);
      if (!why.empty())
        return set_error(DPC_ERR_INTERNAL, "dpc_tree_plan_create: tile run of passes %zu-%zu: %s", ts.l, ts.r - 1,
                         why.c_str());
      st.n_tiles = static_cast<uint32_t>(ts.run.tiles());
      st.ring = ts.run.ring;
      plan->steps.push_back(st);  // owns the buffers from here on
      auto& ps = plan->steps.back();
      // kMaxSegs pieces per tile (check_run bounds the count), so a workgroup finds its tile directly
      std::vector<Seg> padded(static_cast<size_t>(st.n_tiles) * kMaxSegs, Seg{});
      for (size_t t = 0; t < st.n_tiles; ++t)
        std::copy(ts.run.segs.begin() + ts.run.tile_seg[t], ts.run.segs.begin() + ts.run.tile_seg[t + 1],
                  padded.begin() + t * kMaxSegs);
      DPC_HIP_TRY(hipMalloc(&ps.d_segs, padded.size() * sizeof(Seg)), "dpc_tree_plan_create: tiles");
      DPC_HIP_TRY(hipMemcpy(ps.d_segs, padded.data(), padded.size() * sizeof(Seg), hipMemcpyHostToDevice),
                  "dpc_tree_plan_create: tiles");
      plan->tiles += st.n_tiles;
      continue;
    }
    plan->steps.push_back(st);
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
  for (const auto& st : plan->steps) runs += st.kind == tileplan::Step::kTiles ? 1 : 0;
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
      case tileplan::Step::kSmall:
        if (d_dirty_local)
          hipLaunchKernelGGL(tree_small_levels_kernel<true>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld,
                             plan->d_parent, st.run, d_dirty_local, d_dirty_world);
        else
          hipLaunchKernelGGL(tree_small_levels_kernel<false>, dim3(1), dim3(kSmallThreads), 0, s, L, W, ld,
                             plan->d_parent, st.run, d_dirty_local, d_dirty_world);
        break;
      case tileplan::Step::kLevel:
        hipLaunchKernelGGL(tree_level_kernel, dim3(static_cast<unsigned>((st.s1 - st.s0 + kTreeThreads - 1) / kTreeThreads)),
                           dim3(kTreeThreads), 0, s, L, W, ld, plan->d_parent, st.s0, st.s1, d_dirty_local,
                           d_dirty_world);
        break;
      case tileplan::Step::kTiles:
        if (d_dirty_local)
          launch_tiles<true>(*plan, st, L, W, d_dirty_local, d_dirty_world, s);
        else
          launch_tiles<false>(*plan, st, L, W, d_dirty_local, d_dirty_world, s);
        break;
    }
    if (int rc = check_launch("dpc_tree_propagate_plan")) return rc;
  }
  return DPC_OK;
}

}  // extern "C"

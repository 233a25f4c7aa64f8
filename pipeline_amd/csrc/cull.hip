// cull.hip -- single-frustum culling, layout conversion, matrix deltas,
// group bounding box and result diff.
//
// Reference path: dp::culling::cpu::ManagerImpl::cull
// (dp/culling/cpu/src/ManagerImpl.cpp:467-518), GroupCPU::updateOBBs (:114-162),
// ManagerBitSet::calculateBoundingBox (dp/culling/src/ManagerBitSet.cpp:268-306)
// and ResultBitSet::updateChanged (dp/culling/src/ResultBitSet.cpp:61-108).
#include <atomic>
#include <cfloat>
#include <mutex>

#include "cull_math.hpp"
#include "internal.hpp"
#include "batch_types.hpp"

#pragma clang fp contract(off)

namespace dpc {
namespace {

constexpr unsigned kCullThreads = 256;  // persistent streaming form
constexpr unsigned kIndexedThreads = 512;  // plain form with an index array
constexpr unsigned kPlainThreads = 1024;  // plain form without an index array

// With Diff (the cull of Manager::cull, which the result diff follows), the
// wave that stores mask word w also stores the number of bits in which it
// differs from old[w] (bits below n only) as count[w], one byte per word:
// the write pass of the diff then needs no counting launch of its own.
__device__ __forceinline__ uint8_t changed_in_word(unsigned long long old_bits, unsigned long long new_bits, size_t w,
                                                   size_t n)
{
  unsigned long long x = old_bits ^ new_bits;
  const size_t tail = n & 63;
  if (tail && w == (n >> 6)) x &= (~0ull) >> (64 - tail);
  return static_cast<uint8_t>(__popcll(x));
}

// Plain form: one object per lane, one 64-object tile per wave, a grid that
// covers all tiles.  The object's four matrix rows come from four row planes,
// so every load of the wave is one contiguous 1 KiB (rows, box_a) or 512 B
// (box_b) segment; the visibility bits leave as one ballot word per wave.
// Without an index array every box and matrix is read once per launch, so
// they are loaded as streaming data in 1024-thread blocks; with one (matrices
// shared by several objects, e.g. the leaves of a transform tree), plain cached
// loads in 512-thread blocks measured fastest (19.3 us against 19.5-19.7 us with
// 256 threads for the 2.37M objects of a 1M-node tree).
template <bool Indexed, bool Diff>
__global__ __launch_bounds__(Indexed ? kIndexedThreads : kPlainThreads) void cull_kernel(const float4* __restrict__ box_a,
                                                            const float2* __restrict__ box_b,
                                                            const uint32_t* __restrict__ xform, size_t n,
                                                            const float4* __restrict__ planes, size_t ld, Mat16 vp,
                                                            unsigned long long* __restrict__ mask,
                                                            const unsigned long long* __restrict__ old,
                                                            uint8_t* __restrict__ count)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * (Indexed ? kIndexedThreads : kPlainThreads) + threadIdx.x;
  const bool writer = (threadIdx.x & (kWave - 1)) == 0 && i < n;  // lane 0 of a wave holding objects
  const unsigned long long old_bits = Diff && writer ? old[i >> 6] : 0ull;  // loaded with the objects
  const Slot s = make_slot(i, n, Indexed ? xform : nullptr, ld);
  const RawObj r = Indexed ? load_raw(box_a, box_b, planes, ld, s.i, s.m)
                           : load_raw_stream<true>(box_a, box_b, planes, ld, s.i, s.m);
  const bool vis = slot_visible(s, obb_visible(obb_of(r), vp.m));
  const unsigned long long bits = __ballot(vis);
  if (writer) {
    mask[i >> 6] = bits;
    if (Diff) count[i >> 6] = changed_in_word(old_bits, bits, i >> 6, n);
  }
}

// Streaming form: a persistent grid in which every wave walks tiles t, t+W,
// t+2W, ... with two register sets: while one tile is tested, the loads of the
// next one are in flight.  Every wave runs the same number of steps (tiles past
// the end load clamped indices and store nothing), so all loads are
// unconditional and the wait before each compute covers only that tile's loads.
template <bool Indexed, bool Diff>
__device__ __forceinline__ void stream_fetch(const float4* __restrict__ box_a, const float2* __restrict__ box_b,
                                             const uint32_t* __restrict__ xform, size_t n,
                                             const float4* __restrict__ planes, size_t ld, size_t tile, unsigned lane,
                                             const unsigned long long* __restrict__ old, Slot& s, RawObj& r,
                                             unsigned long long& old_bits)
{
  s = make_slot(tile * kWave + lane, n, Indexed ? xform : nullptr, ld);
  r = load_raw(box_a, box_b, planes, ld, s.i, s.m);
  old_bits = Diff && lane == 0 && tile * kWave < n ? old[tile] : 0ull;
  // keep the loads issued here, ahead of the test of the previous tile (the
  // scheduler would otherwise sink them below it to shorten register lifetimes)
  __builtin_amdgcn_sched_barrier(0);
}

template <bool Diff>
__device__ __forceinline__ void stream_test(const Slot& s, const RawObj& r, const Mat16& vp, size_t tile, size_t n,
                                            size_t n_tiles, unsigned lane, unsigned long long* __restrict__ mask,
                                            unsigned long long old_bits, uint8_t* __restrict__ count)
{
  const bool vis = slot_visible(s, obb_visible(obb_of(r), vp.m));
  const unsigned long long bits = __ballot(vis);
  if (lane == 0 && tile < n_tiles) {
    mask[tile] = bits;
    if (Diff) count[tile] = changed_in_word(old_bits, bits, tile, n);
  }
}

template <bool Indexed, bool Diff>
__global__ __launch_bounds__(kCullThreads) void cull_stream_kernel(const float4* __restrict__ box_a,
                                                                   const float2* __restrict__ box_b,
                                                                   const uint32_t* __restrict__ xform, size_t n,
                                                                   const float4* __restrict__ planes, size_t ld,
                                                                   Mat16 vp, unsigned long long* __restrict__ mask,
                                                                   const unsigned long long* __restrict__ old,
                                                                   uint8_t* __restrict__ count)
{
  const unsigned lane = threadIdx.x & (kWave - 1);
  const size_t waves_per_block = blockDim.x / kWave;
  const size_t W = static_cast<size_t>(gridDim.x) * waves_per_block;
  const size_t n_tiles = (n + kWave - 1) / kWave;
  const size_t steps = (n_tiles + W - 1) / W;  // the same for every wave
  size_t t = static_cast<size_t>(blockIdx.x) * waves_per_block + threadIdx.x / kWave;
  Slot sa, sb;
  RawObj ra, rb;
  unsigned long long oa, ob;
  stream_fetch<Indexed, Diff>(box_a, box_b, xform, n, planes, ld, t, lane, old, sa, ra, oa);
  for (size_t k = 0; k < steps; k += 2) {
    stream_fetch<Indexed, Diff>(box_a, box_b, xform, n, planes, ld, t + W, lane, old, sb, rb, ob);
    stream_test<Diff>(sa, ra, vp, t, n, n_tiles, lane, mask, oa, count);
    if (k + 1 >= steps) break;
    stream_fetch<Indexed, Diff>(box_a, box_b, xform, n, planes, ld, t + 2 * W, lane, old, sa, ra, oa);
    stream_test<Diff>(sb, rb, vp, t + W, n, n_tiles, lane, mask, ob, count);
    t += 2 * W;
  }
}

// Row planes <- AoS matrices with a byte stride.  One thread per (matrix, row).
__global__ void from_aos_kernel(const char* __restrict__ aos, size_t stride, size_t count, float4* __restrict__ planes,
                                size_t ld)
{
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * 4) return;
  const size_t k = t >> 2, r = t & 3;
  const float* src = reinterpret_cast<const float*>(aos + k * stride) + 4 * r;
  planes[r * ld + k] = make_float4(src[0], src[1], src[2], src[3]);
}

__global__ void to_aos_kernel(const float4* __restrict__ planes, size_t ld, const uint32_t* __restrict__ slot,
                              size_t count, char* __restrict__ aos, size_t stride)
{
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * 4) return;
  const size_t k = t >> 2, r = t & 3;
  const size_t s = slot ? static_cast<size_t>(slot[k]) : k;
  float* dst = reinterpret_cast<float*>(aos + k * stride) + 4 * r;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (s < ld) v = planes[r * ld + s];
  dst[0] = v.x;
  dst[1] = v.y;
  dst[2] = v.z;
  dst[3] = v.w;
}

__global__ void scatter_kernel(float4* __restrict__ planes, size_t ld, const uint32_t* __restrict__ idx,
                               const float4* __restrict__ src, size_t count)
{
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * 4) return;
  const size_t k = t >> 2, r = t & 3;
  const size_t s = idx[k];
  if (s < ld) planes[r * ld + s] = src[k * 4 + r];
}

// The same scatter with the delta read straight from pinned host memory:
// system-coherent 16-byte buffer loads (sc0 sc1: they see the host's writes
// and are not kept in device caches), so no copy precedes it.  A small grid
// strides over the rows: enough loads in flight to fill PCIe, few waves
// holding CUs while they wait on it.
constexpr int kSystemCoherent = 17;  // buffer load cache policy: sc0 | sc1
constexpr int kRawBuffer = 0x00020000;
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void scatter_host_kernel(float4* __restrict__ planes, uint32_t ld, const uint32_t* idx,
                                                           const float* src, uint32_t count)
{
  const __amdgpu_buffer_rsrc_t ri =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(idx), 0, count * 4u, kRawBuffer);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(src), 0, count * 64u, kRawBuffer);
  // the four row planes as one buffer (ld <= kMaxHostScatterLd keeps it below 4 GB)
  const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(planes, 0, ld * 64u, kRawBuffer);
  const uint32_t rows = count * 4u;
  for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < rows; t += gridDim.x * 256u) {
    const uint32_t k = t >> 2, r = t & 3u;
    // both loads issue before either is waited for; the store is issued on
    // every path (an index outside the planes gets an offset the hardware drops)
    const uint32_t s = __builtin_amdgcn_raw_buffer_load_b32(ri, k * 4u, 0, kSystemCoherent);
    const u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(rs, t * 16u, 0, kSystemCoherent);
    __builtin_amdgcn_raw_buffer_store_b128(v, rp, s < ld ? (r * ld + s) * 16u : 0xFFFFFFF0u, 0, 0);
  }
}

// The scatter of a packed delta (launch_scatter_affine): one thread per matrix,
// three 16-byte loads, the last column from tails or 0 0 0 1 (the bits of
// 0.0f, 0.0f, 0.0f, 1.0f), four 16-byte stores, one to each row plane.
__global__ __launch_bounds__(256) void scatter_affine_kernel(uint4* __restrict__ planes, size_t ld,
                                                             const uint32_t* __restrict__ idx,
                                                             const uint4* __restrict__ packed,
                                                             const uint4* __restrict__ tails, size_t count)
{
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const size_t s = idx[k];
  const uint4 a = packed[3 * k], b = packed[3 * k + 1], c = packed[3 * k + 2];
  const uint4 w = tails ? tails[k] : make_uint4(0u, 0u, 0u, 0x3f800000u);
  if (s >= ld) return;
  planes[s] = make_uint4(a.x, a.y, a.z, w.x);
  planes[ld + s] = make_uint4(b.x, b.y, b.z, w.y);
  planes[2 * ld + s] = make_uint4(c.x, c.y, c.z, w.z);
  planes[3 * ld + s] = make_uint4(a.w, b.w, c.w, w.w);
}

__global__ void gather_u32_kernel(const uint32_t* __restrict__ table, size_t table_len,
                                  const uint32_t* __restrict__ idx, size_t n, uint32_t* __restrict__ out)
{
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t i = idx[k];
  out[k] = i < table_len ? table[i] : 0xFFFFFFFFu;
}

__global__ void boxes_kernel(const float* __restrict__ lu, size_t n, float4* __restrict__ box_a,
                             float2* __restrict__ box_b)
{
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* b = lu + 6 * i;
  // ManagerBitSet.cpp:102-103: lowerLeft = (lower, 1), extent = (upper - lower, 0).
  box_a[i] = make_float4(b[0], b[1], b[2], b[3] - b[0]);
  box_b[i] = make_float2(b[4] - b[1], b[5] - b[2]);
}

// ---- group bounding box -----------------------------------------------------
// Box4f::update (dp/math/Boxnt.h:245-258) keeps the running value unless the new
// point is strictly smaller (lower) / larger (upper).  That is "leftmost extremum"
// selection, which is associative over ordered ranges: combining a range with the
// range that follows it, left operand first, reproduces the sequential scan bit for
// bit (including which of +0/-0 survives and that NaN never enters).
struct __align__(8) Box6 {  // (8-byte aligned: a partial box loads as three dwordx2)
  float lo[3], hi[3];
};

__device__ __forceinline__ void box_take(Box6& acc, const float4& p)
{
  if (acc.lo[0] > p.x) acc.lo[0] = p.x;
  if (acc.hi[0] < p.x) acc.hi[0] = p.x;
  if (acc.lo[1] > p.y) acc.lo[1] = p.y;
  if (acc.hi[1] < p.y) acc.hi[1] = p.y;
  if (acc.lo[2] > p.z) acc.lo[2] = p.z;
  if (acc.hi[2] < p.z) acc.hi[2] = p.z;
}

__device__ __forceinline__ Box6 box_empty()
{
  Box6 b;
  for (int c = 0; c < 3; ++c) {
    b.lo[c] = FLT_MAX;
    b.hi[c] = -FLT_MAX;
  }
  return b;
}

constexpr unsigned kBoxThreads = 256;
constexpr unsigned kBoxWavesPerBlock = kBoxThreads / kWave;
constexpr unsigned kBoxWavesPerCu = 16;

// A running extremum with the index of the object that set it.  Folding
// objects in index order with strict comparisons keeps the first of equal
// values (the reference's sequential Box4f::update, Boxnt.h:245-258); across
// lanes, whose objects interleave, equal values are decided by the lower index.
// Only the sign of a zero extremum depends on this (+0 and -0 compare equal).
struct Box6i {
  float lo[3], hi[3];
  uint32_t ilo[3], ihi[3];
};

__device__ __forceinline__ Box6i box6i_empty()
{
  Box6i b;
  for (int c = 0; c < 3; ++c) {
    b.lo[c] = FLT_MAX;
    b.hi[c] = -FLT_MAX;
    b.ilo[c] = b.ihi[c] = 0xFFFFFFFFu;
  }
  return b;
}

// The smaller value, or the lower index on equal values, of (v, i) and (w, j)
// (selects, no branches: the reductions below are latency-bound).
__device__ __forceinline__ void take_lower(float& v, uint32_t& i, float w, uint32_t j)
{
  const bool take = (v > w) | ((v == w) & (j < i));
  v = take ? w : v;
  i = take ? j : i;
}

__device__ __forceinline__ void take_higher(float& v, uint32_t& i, float w, uint32_t j)
{
  const bool take = (v < w) | ((v == w) & (j < i));
  v = take ? w : v;
  i = take ? j : i;
}

// Wave reduction of per-lane boxes (lane 0 receives the result).  A lane whose
// partner lies beyond the wave reads its own values back, which changes nothing.
__device__ __forceinline__ Box6i wave_box6i_reduce(Box6i b)
{
  const unsigned lane = threadIdx.x & (kWave - 1);
  for (unsigned off = 32; off >= 1; off >>= 1) {
    const int src = static_cast<int>(lane + off < kWave ? lane + off : lane) * 4;
    float lo[3], hi[3];
    uint32_t ilo[3], ihi[3];
    for (int c = 0; c < 3; ++c) {
      lo[c] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(b.lo[c])));
      hi[c] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(b.hi[c])));
      ilo[c] = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(b.ilo[c])));
      ihi[c] = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(b.ihi[c])));
    }
    for (int c = 0; c < 3; ++c) {
      take_lower(b.lo[c], b.ilo[c], lo[c], ilo[c]);
      take_higher(b.hi[c], b.ihi[c], hi[c], ihi[c]);
    }
  }
  return b;
}

__device__ __forceinline__ Box6 box6_of(const Box6i& b)
{
  Box6 r;
  for (int c = 0; c < 3; ++c) {
    r.lo[c] = b.lo[c];
    r.hi[c] = b.hi[c];
  }
  return r;
}

// One object's inputs to the bounding box: its box and its matrix's rows (row
// index clamped to 0 when it lies outside the array; such an object is skipped).
struct BoxIn {
  float4 a;
  float2 b;
  float4 r0, r1, r2, r3;
};

__device__ __forceinline__ BoxIn load_box_in(const float4* __restrict__ box_a, const float2* __restrict__ box_b,
                                             const float4* __restrict__ planes, size_t ld, size_t i, size_t m)
{
  const size_t c = m < ld ? m : 0;
  BoxIn d;
  d.a = box_a[i];
  d.b = box_b[i];
  d.r0 = planes[c];
  d.r1 = planes[ld + c];
  d.r2 = planes[2 * ld + c];
  d.r3 = planes[3 * ld + c];
  return d;
}

// Each wave scans one contiguous chunk of objects, lane l taking objects
// begin + l, begin + l + 64, ... (coalesced loads).  A lane first takes the box
// of its object's eight corners, then folds it into its running box with the
// object's index; one wave reduction per chunk gives the chunk's box.  The
// workgroup's waves hold consecutive chunks, so folding their boxes in wave
// order with strict comparisons keeps the first index among equal extrema; one
// partial per workgroup is stored, in index order of the chunks.
template <bool Indexed>
__global__ __launch_bounds__(kBoxThreads) void bbox_partial_kernel(const float4* __restrict__ box_a,
                                                                   const float2* __restrict__ box_b,
                                                                   const uint32_t* __restrict__ xform, size_t n,
                                                                   const float4* __restrict__ planes, size_t ld,
                                                                   size_t chunk, Box6* __restrict__ partial)
{
  // wave-uniform in scalar registers, so the loop below branches on scalars
  const size_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * kBoxWavesPerBlock + threadIdx.x / kWave);
  const unsigned lane = threadIdx.x & (kWave - 1);
  const size_t begin = wave * chunk;
  size_t end = begin + chunk;
  if (end > n) end = n;
  Box6i acc = box6i_empty();
  // Round k takes object begin + lane + 64k.  The loads are clamped and
  // unconditional and run ahead of the folds, in two register sets that take
  // turns (no copies, which would wait for the loads): the box and matrix rows
  // of round k + 1 and the matrix index of round k + 2 are in flight while round
  // k is folded.
  const size_t count = end > begin ? end - begin : 0;
  const size_t rounds = (count + kWave - 1) / kWave;  // the same on every lane of the wave
  auto object_of = [&](size_t k) -> size_t {
    const size_t i = begin + lane + k * kWave;
    return i < end ? i : begin;  // (begin < end whenever rounds > 0)
  };
  auto matrix_of = [&](size_t i) -> size_t { return Indexed ? static_cast<size_t>(xform[i]) : i; };
  // Folds round k's object (index i, matrix m, inputs d) into the lane's box.
  auto fold = [&](const BoxIn& d, size_t i, size_t m, size_t k) {
    // a lane past the chunk's end, or an object whose matrix index is outside
    // the array, adds nothing
    const bool valid = (begin + lane + k * kWave < end) & (m < ld);
    // ManagerBitSet.cpp:282-294: v0 = lower * M, x/y/z = extent[k] * row k, same add tree as the cull.
    const Obb o = make_obb(d.a, d.b, d.r0, d.r1, d.r2, d.r3);
    const float4 v1 = add4(o.p, o.ex);
    const float4 v2 = add4(o.p, o.ey);
    const float4 v3 = add4(v1, o.ey);
    const float4 v4 = add4(o.p, o.ez);
    const float4 v5 = add4(v1, o.ez);
    const float4 v6 = add4(v2, o.ez);
    const float4 v7 = add4(v3, o.ez);
    Box6 mine = box_empty();
    box_take(mine, o.p);
    box_take(mine, v1);
    box_take(mine, v2);
    box_take(mine, v3);
    box_take(mine, v4);
    box_take(mine, v5);
    box_take(mine, v6);
    box_take(mine, v7);
    const uint32_t idx = static_cast<uint32_t>(i - begin);
    for (int c = 0; c < 3; ++c) {
      const bool lower = valid & (acc.lo[c] > mine.lo[c]);
      acc.lo[c] = lower ? mine.lo[c] : acc.lo[c];
      acc.ilo[c] = lower ? idx : acc.ilo[c];
      const bool higher = valid & (acc.hi[c] < mine.hi[c]);
      acc.hi[c] = higher ? mine.hi[c] : acc.hi[c];
      acc.ihi[c] = higher ? idx : acc.ihi[c];
    }
  };
  if (rounds) {
    size_t ia = object_of(0), ma = matrix_of(ia);
    BoxIn a = load_box_in(box_a, box_b, planes, ld, ia, ma);
    size_t in = object_of(1), mn = matrix_of(in);  // the next round's object and matrix index
    BoxIn b;
    size_t ib = 0, mb = 0;
    for (size_t k = 0; k < rounds; k += 2) {
      // round k is in a; load round k + 1 into b and the index of round k + 2
      ib = in;
      mb = mn;
      in = object_of(k + 2);
      mn = matrix_of(in);
      b = load_box_in(box_a, box_b, planes, ld, ib, mb);
      __builtin_amdgcn_sched_barrier(0);  // keep these loads ahead of the fold
      fold(a, ia, ma, k);
      if (k + 1 >= rounds) break;
      // round k + 1 is in b; load round k + 2 into a and the index of round k + 3
      ia = in;
      ma = mn;
      in = object_of(k + 3);
      mn = matrix_of(in);
      a = load_box_in(box_a, box_b, planes, ld, ia, ma);
      __builtin_amdgcn_sched_barrier(0);
      fold(b, ib, mb, k + 1);
    }
  }
  __shared__ Box6 waves[kBoxWavesPerBlock];
  const Box6 box = box6_of(wave_box6i_reduce(acc));
  if (lane == 0) waves[threadIdx.x / kWave] = box;
  __syncthreads();
  if (threadIdx.x != 0) return;
  Box6 all = waves[0];
  for (unsigned w = 1; w < kBoxWavesPerBlock; ++w)
    for (int c = 0; c < 3; ++c) {
      if (waves[w].lo[c] < all.lo[c]) all.lo[c] = waves[w].lo[c];
      if (waves[w].hi[c] > all.hi[c]) all.hi[c] = waves[w].hi[c];
    }
  partial[blockIdx.x] = all;
}

constexpr unsigned kBoxFinalThreads = 256;

// One workgroup folds the partial boxes.  Thread t takes partials t, t + 256,
// ... (coalesced loads, four per round issued together) in order, keeping the
// index of the partial that set each extremum; waves and then the workgroup's
// four wave boxes are reduced by value and then partial index (partials are in
// index order of their chunks).
__global__ __launch_bounds__(kBoxFinalThreads) void bbox_final_kernel(const Box6* __restrict__ partial,
                                                                      size_t n_partial, float* __restrict__ out6)
{
  __shared__ Box6i waves[kBoxFinalThreads / kWave];
  const unsigned t = threadIdx.x, lane = t & (kWave - 1);
  Box6i run = box6i_empty();
  // four partials per thread and round, all loaded before any is folded (a
  // clamped index past the end loads the last partial again, which is skipped)
  constexpr unsigned kU = 4;
  for (size_t k0 = t; k0 < n_partial; k0 += kU * kBoxFinalThreads) {
    Box6 b[kU];
#pragma unroll
    for (unsigned u = 0; u < kU; ++u) {
      const size_t k = k0 + u * kBoxFinalThreads;
      b[u] = partial[k < n_partial ? k : n_partial - 1];
    }
#pragma unroll
    for (unsigned u = 0; u < kU; ++u) {
      // a slot past the end folds the empty box (it can change no value)
      const size_t k = k0 + u * kBoxFinalThreads;
      const bool valid = k < n_partial;
      for (int c = 0; c < 3; ++c) {
        take_lower(run.lo[c], run.ilo[c], valid ? b[u].lo[c] : FLT_MAX, static_cast<uint32_t>(k));
        take_higher(run.hi[c], run.ihi[c], valid ? b[u].hi[c] : -FLT_MAX, static_cast<uint32_t>(k));
      }
    }
  }
  const Box6i wb = wave_box6i_reduce(run);
  if (lane == 0) waves[t / kWave] = wb;
  __syncthreads();
  if (t != 0) return;
  Box6i all = waves[0];
  for (unsigned w = 1; w < kBoxFinalThreads / kWave; ++w)
    for (int c = 0; c < 3; ++c) {
      take_lower(all.lo[c], all.ilo[c], waves[w].lo[c], waves[w].ilo[c]);
      take_higher(all.hi[c], all.ihi[c], waves[w].hi[c], waves[w].ihi[c]);
    }
  const Box6 acc = box6_of(all);
  {
    // return Box3f(lower.xyz, upper.xyz): Boxnt(p0, p1) runs init(), update(p0), update(p1)
    // (dp/math/Boxnt.h:213-219), which turns the empty box into (-FLT_MAX, +FLT_MAX).
    Box6 r = box_empty();
    box_take(r, make_float4(acc.lo[0], acc.lo[1], acc.lo[2], 0.f));
    box_take(r, make_float4(acc.hi[0], acc.hi[1], acc.hi[2], 0.f));
    for (int c = 0; c < 3; ++c) {
      out6[c] = r.lo[c];
      out6[3 + c] = r.hi[c];
    }
  }
}

size_t bbox_waves(size_t n)
{
  // kBoxWavesPerCu waves per CU of partials at most, at least 64 objects each.
  size_t waves = (n + kWave - 1) / kWave;
  const size_t cap = static_cast<size_t>(device_cu_count()) * kBoxWavesPerCu;
  if (waves > cap) waves = cap;
  if (waves == 0) waves = 1;
  waves = (waves + kBoxWavesPerBlock - 1) / kBoxWavesPerBlock * kBoxWavesPerBlock;
  return waves;
}

// ---- result diff --------------------------------------------------------------
constexpr unsigned kDiffThreads = 256;  // one 64-bit word per thread

__device__ __forceinline__ unsigned long long diff_word(const unsigned long long* old_w,
                                                        const unsigned long long* new_w, size_t w, size_t n_words,
                                                        size_t n_bits)
{
  if (w >= n_words) return 0ull;
  unsigned long long x = old_w[w] ^ new_w[w];
  const size_t tail = n_bits & 63;
  if (w == n_words - 1 && tail) x &= (~0ull) >> (64 - tail);
  return x;
}

// diff_word without a branch: both words are loaded (from a clamped index)
// on every lane, so the loads of a kernel's later steps can be issued before
// anything waits for these.
__device__ __forceinline__ unsigned long long diff_word_clamped(const unsigned long long* old_w,
                                                                const unsigned long long* new_w, size_t w,
                                                                size_t n_words, size_t n_bits,
                                                                unsigned long long& new_bits)
{
  const size_t c = w < n_words ? w : n_words - 1;  // n_words >= 1 wherever a diff kernel runs
  new_bits = new_w[c];
  unsigned long long x = old_w[c] ^ new_bits;
  const size_t tail = n_bits & 63;
  const unsigned long long keep = (tail && c == n_words - 1) ? (~0ull) >> (64 - tail) : ~0ull;
  return w < n_words ? (x & keep) : 0ull;
}

// Exclusive scan over a block of Threads threads (lds: Threads / 64 + 1 words).
template <unsigned Threads = kDiffThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t* total)
{
  // Wave-level inclusive scan, then scan of the wave totals.
  const unsigned lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t x = v;
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= static_cast<unsigned>(off)) x += y;
  }
  if (lane == kWave - 1) lds[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (unsigned k = 0; k < Threads / kWave; ++k) {
      const uint32_t t = lds[k];
      lds[k] = run;
      run += t;
    }
    lds[Threads / kWave] = run;
  }
  __syncthreads();
  const uint32_t r = lds[wave] + x - v;
  *total = lds[Threads / kWave];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kDiffThreads) void diff_count_kernel(const unsigned long long* __restrict__ old_w,
                                                                  const unsigned long long* __restrict__ new_w,
                                                                  size_t n_words, size_t n_bits,
                                                                  uint32_t* __restrict__ block_count)
{
  __shared__ uint32_t lds[kDiffThreads / kWave + 1];
  const size_t w = static_cast<size_t>(blockIdx.x) * kDiffThreads + threadIdx.x;
  const uint32_t c = __popcll(diff_word(old_w, new_w, w, n_words, n_bits));
  uint32_t total;
  block_exclusive_scan(c, lds, &total);
  if (threadIdx.x == 0) block_count[blockIdx.x] = total;
}

// Single block: exclusive scan of block counts in place; writes the grand total.
__global__ __launch_bounds__(kDiffThreads) void diff_scan_kernel(uint32_t* __restrict__ block_count, size_t n_blocks,
                                                                 uint32_t* __restrict__ d_count)
{
  __shared__ uint32_t lds[kDiffThreads / kWave + 1];
  uint32_t carry = 0;
  for (size_t base = 0; base < n_blocks; base += kDiffThreads) {
    const size_t k = base + threadIdx.x;
    const uint32_t v = k < n_blocks ? block_count[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, lds, &total);
    if (k < n_blocks) block_count[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *d_count = carry;
}

__global__ __launch_bounds__(kDiffThreads) void diff_write_kernel(unsigned long long* __restrict__ old_w,
                                                                  const unsigned long long* __restrict__ new_w,
                                                                  size_t n_words, size_t n_bits,
                                                                  const uint32_t* __restrict__ block_offset,
                                                                  uint32_t* __restrict__ out, int update_old)
{
  __shared__ uint32_t lds[kDiffThreads / kWave + 1];
  const size_t w = static_cast<size_t>(blockIdx.x) * kDiffThreads + threadIdx.x;
  unsigned long long x = diff_word(old_w, new_w, w, n_words, n_bits);
  uint32_t total;
  uint32_t pos = block_offset[blockIdx.x] + block_exclusive_scan(__popcll(x), lds, &total);
  while (x) {
    const unsigned b = __ffsll(static_cast<long long>(x)) - 1;
    out[pos++] = static_cast<uint32_t>(w * 64 + b);
    x &= x - 1;
  }
  if (update_old && w < n_words) old_w[w] = new_w[w];
}

// The write pass for at most kDiffDirectBlocks blocks: each block sums the
// counts of the blocks before it itself (no separate scan launch); the last
// block writes the grand total.
constexpr size_t kDiffDirectBlocks = 1024;

__global__ __launch_bounds__(kDiffThreads) void diff_write_direct_kernel(unsigned long long* __restrict__ old_w,
                                                                         const unsigned long long* __restrict__ new_w,
                                                                         size_t n_words, size_t n_bits,
                                                                         const uint32_t* __restrict__ block_count,
                                                                         uint32_t* __restrict__ out, int update_old,
                                                                         uint32_t* __restrict__ d_count)
{
  __shared__ uint32_t lds[kDiffThreads / kWave + 1];
  // this thread's words first: their loads are then in flight with those of the counts
  const size_t w = static_cast<size_t>(blockIdx.x) * kDiffThreads + threadIdx.x;
  unsigned long long new_bits;
  unsigned long long x = diff_word_clamped(old_w, new_w, w, n_words, n_bits, new_bits);
  // this thread's share of the counts of blocks [0, blockIdx.x): clamped loads, all issued at once
  constexpr unsigned kCountsPerThread = kDiffDirectBlocks / kDiffThreads;
  uint32_t v[kCountsPerThread];
#pragma unroll
  for (unsigned q = 0; q < kCountsPerThread; ++q) {
    const unsigned k = threadIdx.x + q * kDiffThreads;
    v[q] = block_count[k < blockIdx.x ? k : 0u];
  }
  uint32_t before = 0;
#pragma unroll
  for (unsigned q = 0; q < kCountsPerThread; ++q) before += threadIdx.x + q * kDiffThreads < blockIdx.x ? v[q] : 0u;
  uint32_t offset;
  block_exclusive_scan(before, lds, &offset);
  uint32_t total;
  uint32_t pos = offset + block_exclusive_scan(__popcll(x), lds, &total);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_count = offset + total;
  while (x) {
    const unsigned b = __ffsll(static_cast<long long>(x)) - 1;
    out[pos++] = static_cast<uint32_t>(w * 64 + b);
    x &= x - 1;
  }
  if (update_old && w < n_words) old_w[w] = new_bits;
}

// The write pass after a cull that stored one change count per mask word
// (cull_kernel / cull_stream_kernel with Diff): each block of kDiffFusedThreads
// words adds up the counts of the words before its own, sixteen bytes per
// thread and load, and then writes as diff_write_direct_kernel.  For at most
// kDiffFusedWords mask words (2^22 objects): every thread then needs at most
// kDiffFusedWords / 16 / kDiffFusedThreads = 4 loads of counts, issued at once.
constexpr unsigned kDiffFusedThreads = 1024;
constexpr size_t kDiffFusedWords = 65536;

__device__ __forceinline__ uint32_t byte_sum(uint32_t v, uint32_t acc) { return __builtin_amdgcn_sad_u8(v, 0u, acc); }

__global__ __launch_bounds__(kDiffFusedThreads) void diff_write_fused_kernel(unsigned long long* __restrict__ old_w,
                                                                             const unsigned long long* __restrict__ new_w,
                                                                             size_t n_words, size_t n_bits,
                                                                             const uint4* __restrict__ word_count,
                                                                             uint32_t* __restrict__ out, int update_old,
                                                                             uint32_t* __restrict__ d_count)
{
  __shared__ uint32_t lds[kDiffFusedThreads / kWave + 1];
  // this thread's words first: their loads are then in flight with those of the counts
  const size_t w = static_cast<size_t>(blockIdx.x) * kDiffFusedThreads + threadIdx.x;
  unsigned long long new_bits;
  unsigned long long x = diff_word_clamped(old_w, new_w, w, n_words, n_bits, new_bits);
  // this thread's share of the counts of words [0, kDiffFusedThreads * blockIdx.x), clamped loads
  constexpr unsigned kChunksPerThread = kDiffFusedWords / 16 / kDiffFusedThreads;
  const size_t chunks = static_cast<size_t>(blockIdx.x) * (kDiffFusedThreads / 16);
  uint4 c[kChunksPerThread];
#pragma unroll
  for (unsigned q = 0; q < kChunksPerThread; ++q) {
    const size_t k = threadIdx.x + static_cast<size_t>(q) * kDiffFusedThreads;
    c[q] = word_count[k < chunks ? k : 0];
  }
  uint32_t before = 0;
#pragma unroll
  for (unsigned q = 0; q < kChunksPerThread; ++q) {
    const uint32_t part = byte_sum(c[q].w, byte_sum(c[q].z, byte_sum(c[q].y, byte_sum(c[q].x, 0u))));
    before += threadIdx.x + static_cast<size_t>(q) * kDiffFusedThreads < chunks ? part : 0u;
  }
  uint32_t offset;
  block_exclusive_scan<kDiffFusedThreads>(before, lds, &offset);
  uint32_t total;
  uint32_t pos = offset + block_exclusive_scan<kDiffFusedThreads>(__popcll(x), lds, &total);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_count = offset + total;
  while (x) {
    const unsigned b = __ffsll(static_cast<long long>(x)) - 1;
    if (pos < n_bits) out[pos] = static_cast<uint32_t>(w * 64 + b);  // (always, when the counts match the masks)
    ++pos;
    x &= x - 1;
  }
  if (update_old && w < n_words) old_w[w] = new_bits;
}

unsigned grid_for(size_t work, unsigned per_block)
{
  return static_cast<unsigned>((work + per_block - 1) / per_block);
}

std::atomic<uint32_t> g_cull_waves_per_cu{0};  // the plain form (measured fastest on gfx950)

dpc_cull_config current_cull_config()
{
  dpc_cull_config c;
  c.waves_per_cu = g_cull_waves_per_cu.load(std::memory_order_relaxed);
  return c;
}

// Argument checks of the single-frustum cull (what: the caller's name).
int check_cull_args(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
                    const float* d_planes, size_t ld, const float* vp, const uint64_t* d_mask, const char* what)
{
  if (!d_box_a || !d_box_b || !d_planes || !d_mask || !vp)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: null pointer argument", what);
  if (!aligned(d_box_a, 16) || !aligned(d_planes, 16) || !aligned(d_box_b, 8) || !aligned(d_mask, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: misaligned device pointer", what);
  if (!d_xform && ld < n_objects)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: %zu objects without indices need ld >= n (ld=%zu)", what,
                     n_objects, ld);
  if (ld == 0) return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: no matrices (ld = 0)", what);
  return DPC_OK;
}

// The cull's launch in the tuned shape; with Diff, also the per-word change
// counts against d_old (count: one byte per mask word).
template <bool Diff>
int launch_cull(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
                const float* d_planes, size_t ld, const float* vp, uint64_t* d_mask, const uint64_t* d_old,
                uint8_t* count, void* stream)
{
  Mat16 m;
  for (int k = 0; k < 16; ++k) m.m[k] = vp[k];
  const auto* a = reinterpret_cast<const float4*>(d_box_a);
  const auto* b = reinterpret_cast<const float2*>(d_box_b);
  const auto* p = reinterpret_cast<const float4*>(d_planes);
  auto* out = reinterpret_cast<unsigned long long*>(d_mask);
  const auto* old = reinterpret_cast<const unsigned long long*>(d_old);
  const dpc_cull_config cfg = current_cull_config();
  const size_t tiles = (n_objects + kWave - 1) / kWave;
  hipStream_t s = as_stream(stream);
  if (cfg.waves_per_cu == 0) {
    if (d_xform)
      hipLaunchKernelGGL((cull_kernel<true, Diff>), dim3(grid_for(n_objects, kIndexedThreads)), dim3(kIndexedThreads), 0, s,
                         a, b, d_xform, n_objects, p, ld, m, out, old, count);
    else
      hipLaunchKernelGGL((cull_kernel<false, Diff>), dim3(grid_for(n_objects, kPlainThreads)), dim3(kPlainThreads), 0,
                         s, a, b, d_xform, n_objects, p, ld, m, out, old, count);
  } else {
    const size_t waves_per_block = kCullThreads / kWave;
    size_t blocks = static_cast<size_t>(device_cu_count()) * cfg.waves_per_cu / waves_per_block;
    const size_t needed = (tiles + waves_per_block - 1) / waves_per_block;
    if (blocks > needed) blocks = needed;
    if (blocks == 0) blocks = 1;
    if (d_xform)
      hipLaunchKernelGGL((cull_stream_kernel<true, Diff>), dim3(static_cast<unsigned>(blocks)), dim3(kCullThreads), 0,
                         s, a, b, d_xform, n_objects, p, ld, m, out, old, count);
    else
      hipLaunchKernelGGL((cull_stream_kernel<false, Diff>), dim3(static_cast<unsigned>(blocks)), dim3(kCullThreads),
                         0, s, a, b, d_xform, n_objects, p, ld, m, out, old, count);
  }
  return check_launch("cull_kernel");
}

}  // namespace
}  // namespace dpc

using namespace dpc;

extern "C" {

int dpc_cull(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
             const float* d_planes, size_t ld, const float vp[16], uint64_t* d_mask, void* stream)
{
  if (n_objects == 0) return DPC_OK;
  if (int rc = check_cull_args(d_box_a, d_box_b, d_xform, n_objects, d_planes, ld, vp, d_mask, "dpc_cull")) return rc;
  return launch_cull<false>(d_box_a, d_box_b, d_xform, n_objects, d_planes, ld, vp, d_mask, nullptr, nullptr, stream);
}

int dpc_cull_get_config(dpc_cull_config* cfg)
{
  if (!cfg) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_get_config: null output");
  *cfg = current_cull_config();
  return DPC_OK;
}

int dpc_cull_set_config(const dpc_cull_config* cfg)
{
  if (!cfg || cfg->waves_per_cu > 64)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_cull_set_config: waves_per_cu must be 0..64");
  g_cull_waves_per_cu.store(cfg->waves_per_cu, std::memory_order_relaxed);
  return DPC_OK;
}

int dpc_matrices_from_aos(const void* d_aos, size_t stride_bytes, size_t count, float* d_planes, size_t ld,
                          void* stream)
{
  if (count == 0) return DPC_OK;
  if (!d_aos || !d_planes) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_from_aos: null pointer");
  if (stride_bytes < 64 || stride_bytes % 4 || !aligned(d_aos, 4) || !aligned(d_planes, 16) || ld < count)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_from_aos: bad stride/alignment/ld");
  hipLaunchKernelGGL(from_aos_kernel, dim3(grid_for(count * 4, 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const char*>(d_aos), stride_bytes, count, reinterpret_cast<float4*>(d_planes), ld);
  return check_launch("from_aos_kernel");
}

int dpc_matrices_to_aos(const float* d_planes, size_t ld, const uint32_t* d_slot, size_t count, void* d_aos,
                        size_t stride_bytes, void* stream)
{
  if (count == 0) return DPC_OK;
  if (!d_aos || !d_planes) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_to_aos: null pointer");
  if (stride_bytes < 64 || stride_bytes % 4 || !aligned(d_aos, 4) || !aligned(d_planes, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_to_aos: bad stride/alignment");
  if (!d_slot && ld < count) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_to_aos: ld < count");
  hipLaunchKernelGGL(to_aos_kernel, dim3(grid_for(count * 4, 256)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const float4*>(d_planes), ld, d_slot, count, static_cast<char*>(d_aos),
                     stride_bytes);
  return check_launch("to_aos_kernel");
}

int dpc_matrices_scatter(float* d_planes, size_t ld, const uint32_t* d_idx, const float* d_src, size_t count,
                         void* stream)
{
  if (count == 0) return DPC_OK;
  if (!d_planes || !d_idx || !d_src) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_scatter: null pointer");
  if (!aligned(d_planes, 16) || !aligned(d_src, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_matrices_scatter: misaligned pointer");
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(count * 4, 256)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<float4*>(d_planes), ld, d_idx, reinterpret_cast<const float4*>(d_src), count);
  return check_launch("scatter_kernel");
}

}  // extern "C"

namespace dpc {

int launch_scatter_host(float* d_planes, size_t ld, const uint32_t* idx, const float* src, size_t count,
                        void* stream)
{
  if (count == 0) return DPC_OK;
  if (count > kMaxHostScatter || ld > kMaxHostScatterLd)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "launch_scatter_host: %zu matrices into %zu slots exceed %u into %u",
                     count, ld, kMaxHostScatter, kMaxHostScatterLd);
  // 64 workgroups (256 CUs hold 4x more): ~16k rows of 16 B in flight, well above PCIe's bandwidth-latency product
  const unsigned grid = std::min<unsigned>(grid_for(count * 4, 256), 64u);
  hipLaunchKernelGGL(scatter_host_kernel, dim3(grid), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<float4*>(d_planes), static_cast<uint32_t>(ld), idx, src, static_cast<uint32_t>(count));
  return check_launch("scatter_host_kernel");
}

int launch_scatter_affine(float* d_planes, size_t ld, const uint32_t* d_idx, const void* d_packed,
                          const void* d_tails, size_t count, void* stream)
{
  if (count == 0) return DPC_OK;
  if (!d_planes || !d_idx || !d_packed)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "launch_scatter_affine: null pointer");
  if (!aligned(d_planes, 16) || !aligned(d_packed, 16) || (d_tails && !aligned(d_tails, 16)))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "launch_scatter_affine: misaligned pointer");
  hipLaunchKernelGGL(scatter_affine_kernel, dim3(grid_for(count, 256)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<uint4*>(d_planes), ld, d_idx, static_cast<const uint4*>(d_packed),
                     static_cast<const uint4*>(d_tails), count);
  return check_launch("scatter_affine_kernel");
}

}  // namespace dpc

extern "C" {

int dpc_gather_u32(const uint32_t* d_table, size_t table_len, const uint32_t* d_idx, size_t n, uint32_t* d_out,
                   void* stream)
{
  if (n == 0) return DPC_OK;
  if (!d_idx || !d_out || (table_len && !d_table))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_gather_u32: null pointer");
  hipLaunchKernelGGL(gather_u32_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), d_table, table_len,
                     d_idx, n, d_out);
  return check_launch("gather_u32_kernel");
}

int dpc_boxes_from_lower_upper(const float* d_lower_upper, size_t n, float* d_box_a, float* d_box_b, void* stream)
{
  if (n == 0) return DPC_OK;
  if (!d_lower_upper || !d_box_a || !d_box_b || !aligned(d_box_a, 16) || !aligned(d_box_b, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_boxes_from_lower_upper: bad pointer");
  hipLaunchKernelGGL(boxes_kernel, dim3(grid_for(n, 256)), dim3(256), 0, as_stream(stream), d_lower_upper, n,
                     reinterpret_cast<float4*>(d_box_a), reinterpret_cast<float2*>(d_box_b));
  return check_launch("boxes_kernel");
}

size_t dpc_bbox_workspace(size_t n_objects) { return bbox_waves(n_objects) / kBoxWavesPerBlock * sizeof(Box6); }

int dpc_bbox(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
             const float* d_planes, size_t ld, float* d_out6, void* d_workspace, void* stream)
{
  if (!d_out6 || !d_workspace || !aligned(d_workspace, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_bbox: null output/workspace (or workspace not 8-byte aligned)");
  if (n_objects && (!d_box_a || !d_box_b || !d_planes || !aligned(d_box_a, 16) || !aligned(d_box_b, 8) ||
                    !aligned(d_planes, 16)))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_bbox: bad object/matrix pointer");
  if (n_objects && !d_xform && ld < n_objects)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_bbox: objects without indices need ld >= n");
  const size_t waves = bbox_waves(n_objects);
  size_t chunk = (n_objects + waves - 1) / waves;
  chunk = (chunk + kWave - 1) / kWave * kWave;
  if (chunk == 0) chunk = kWave;
  auto* partial = static_cast<Box6*>(d_workspace);
  const dim3 grid(static_cast<unsigned>(waves / kBoxWavesPerBlock));
  if (d_xform)
    hipLaunchKernelGGL(bbox_partial_kernel<true>, grid, dim3(kBoxThreads), 0, as_stream(stream),
                       reinterpret_cast<const float4*>(d_box_a), reinterpret_cast<const float2*>(d_box_b), d_xform,
                       n_objects, reinterpret_cast<const float4*>(d_planes), ld, chunk, partial);
  else
    hipLaunchKernelGGL(bbox_partial_kernel<false>, grid, dim3(kBoxThreads), 0, as_stream(stream),
                       reinterpret_cast<const float4*>(d_box_a), reinterpret_cast<const float2*>(d_box_b), d_xform,
                       n_objects, reinterpret_cast<const float4*>(d_planes), ld, chunk, partial);
  int rc = check_launch("bbox_partial_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(kBoxFinalThreads), 0, as_stream(stream), partial,
                     waves / kBoxWavesPerBlock, d_out6);
  return check_launch("bbox_final_kernel");
}

size_t dpc_result_diff_workspace(size_t n_bits)
{
  const size_t n_words = (n_bits + 63) / 64;
  const size_t blocks = (n_words + kDiffThreads - 1) / kDiffThreads;
  return (blocks ? blocks : 1) * sizeof(uint32_t);
}

}  // extern "C"

namespace dpc {
namespace {

// Argument checks shared by both result-diff entry points (what: the caller's name).
int check_result_diff_args(const uint64_t* d_old, const uint64_t* d_new, size_t n_bits,
                           const uint32_t* d_changed_idx, const uint32_t* d_count, const char* what)
{
  if (!d_count) return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: null count", what);
  if (n_bits == 0) return DPC_OK;
  if (!d_old || !d_new || !d_changed_idx || !aligned(d_old, 8) || !aligned(d_new, 8))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: bad mask pointer", what);
  if (n_bits > 0xFFFFFFFFull) return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: more than 2^32 bits", what);
  return DPC_OK;
}

// The result diff's launches, with counts = one u32 per block of kDiffThreads
// mask words (every one written by diff_count_kernel before it is read).
int launch_result_diff(uint64_t* d_old, const uint64_t* d_new, size_t n_bits, uint32_t* d_changed_idx,
                       uint32_t* d_count, int update_old, uint32_t* counts, void* stream)
{
  if (n_bits == 0) return dpc_fill_u32(d_count, 0u, 1, stream);
  const size_t n_words = (n_bits + 63) / 64;
  const unsigned blocks = grid_for(n_words, kDiffThreads);
  auto* o = reinterpret_cast<unsigned long long*>(d_old);
  auto* nw = reinterpret_cast<const unsigned long long*>(d_new);
  hipLaunchKernelGGL(diff_count_kernel, dim3(blocks), dim3(kDiffThreads), 0, as_stream(stream), o, nw, n_words, n_bits,
                     counts);
  int rc = check_launch("diff_count_kernel");
  if (rc) return rc;
  if (blocks <= kDiffDirectBlocks) {
    hipLaunchKernelGGL(diff_write_direct_kernel, dim3(blocks), dim3(kDiffThreads), 0, as_stream(stream), o, nw,
                       n_words, n_bits, counts, d_changed_idx, update_old, d_count);
    return check_launch("diff_write_direct_kernel");
  }
  hipLaunchKernelGGL(diff_scan_kernel, dim3(1), dim3(kDiffThreads), 0, as_stream(stream), counts,
                     static_cast<size_t>(blocks), d_count);
  rc = check_launch("diff_scan_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(diff_write_kernel, dim3(blocks), dim3(kDiffThreads), 0, as_stream(stream), o, nw, n_words, n_bits,
                     counts, d_changed_idx, update_old);
  return check_launch("diff_write_kernel");
}

}  // namespace
}  // namespace dpc

// ---- allocation-free result diff ----------------------------------------------
// The per-block counts of dpc_result_diff_noalloc live in this array, one copy
// per device, which the library holds until the process ends (256 KiB: 65,536
// blocks of 256 mask words, i.e. up to 2^30 bits).  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in): it looks up
// each device's address once with hipGetSymbolAddress and passes that.
static_assert(DPC_RESULT_DIFF_NOALLOC_MAX_BITS % (64ull * 256ull) == 0, "whole blocks of 256 mask words");
static __device__ __align__(256) uint32_t scratch[DPC_RESULT_DIFF_NOALLOC_MAX_BITS / (64ull * 256ull)];

namespace dpc {
namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[kScratchDevices] = {};  // per device, looked up on its first call

// The current device's copy of the scratch array (no allocation: a fixed table
// of per-device addresses, filled on each device's first call).
int scratch_of_current_device(uint32_t** ptr, const char* what)
{
  int dev = 0;
  DPC_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kScratchDevices)
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: device %d is beyond the %d devices served", what, dev,
                     kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(DPC_ERR_HIP, "%s: hipGetSymbolAddress returned no address", what);
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return DPC_OK;
}

}  // namespace
}  // namespace dpc

extern "C" {

int dpc_result_diff(uint64_t* d_old, const uint64_t* d_new, size_t n_bits, uint32_t* d_changed_idx, uint32_t* d_count,
                    int update_old, void* d_workspace, void* stream)
{
  if (!d_workspace) return set_error(DPC_ERR_INVALID_ARGUMENT, "dpc_result_diff: null workspace");
  const int rc = check_result_diff_args(d_old, d_new, n_bits, d_changed_idx, d_count, "dpc_result_diff");
  if (rc) return rc;
  return launch_result_diff(d_old, d_new, n_bits, d_changed_idx, d_count, update_old,
                            static_cast<uint32_t*>(d_workspace), stream);
}

int dpc_result_diff_noalloc(uint64_t* d_old, const uint64_t* d_new, size_t n_bits, uint32_t* d_changed_idx,
                            uint32_t* d_count, int update_old, void* stream)
{
  int rc = check_result_diff_args(d_old, d_new, n_bits, d_changed_idx, d_count, "dpc_result_diff_noalloc");
  if (rc) return rc;
  if (n_bits > DPC_RESULT_DIFF_NOALLOC_MAX_BITS)
    return set_error(DPC_ERR_INVALID_ARGUMENT,
                     "dpc_result_diff_noalloc: %zu bits exceed the %llu its fixed scratch serves (use dpc_result_diff)",
                     n_bits, static_cast<unsigned long long>(DPC_RESULT_DIFF_NOALLOC_MAX_BITS));
  uint32_t* counts = nullptr;
  rc = scratch_of_current_device(&counts, "dpc_result_diff_noalloc");
  if (rc) return rc;
  return launch_result_diff(d_old, d_new, n_bits, d_changed_idx, d_count, update_old, counts, stream);
}

size_t dpc_cull_diff_workspace(size_t n_objects)
{
  const size_t words = (n_objects + 63) / 64;
  const size_t counts = (words + 15) / 16 * 16;  // one byte per mask word, read sixteen at a time
  const size_t diff = dpc_result_diff_workspace(n_objects);
  return counts > diff ? counts : diff;
}

int dpc_cull_diff(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
                  const float* d_planes, size_t ld, const float vp[16], uint64_t* d_mask, uint64_t* d_old,
                  uint32_t* d_changed_idx, uint32_t* d_count, int update_old, void* d_workspace, void* stream)
{
  const char* what = "dpc_cull_diff";
  int rc = check_result_diff_args(d_old, d_mask, n_objects, d_changed_idx, d_count, what);
  if (rc) return rc;
  if (n_objects == 0) return dpc_fill_u32(d_count, 0u, 1, stream);
  if ((rc = check_cull_args(d_box_a, d_box_b, d_xform, n_objects, d_planes, ld, vp, d_mask, what))) return rc;
  if (!d_workspace || !aligned(d_workspace, 16))
    return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: null or misaligned workspace", what);
  if (d_old == d_mask) return set_error(DPC_ERR_INVALID_ARGUMENT, "%s: the old and new masks are one buffer", what);
  const size_t n_words = (n_objects + 63) / 64;
  if (n_words > kDiffFusedWords) {  // the prefix sums would grow quadratically: count in a launch of its own
    rc = launch_cull<false>(d_box_a, d_box_b, d_xform, n_objects, d_planes, ld, vp, d_mask, nullptr, nullptr, stream);
    if (rc) return rc;
    return launch_result_diff(d_old, d_mask, n_objects, d_changed_idx, d_count, update_old,
                              static_cast<uint32_t*>(d_workspace), stream);
  }
  auto* counts = static_cast<uint8_t*>(d_workspace);
  rc = launch_cull<true>(d_box_a, d_box_b, d_xform, n_objects, d_planes, ld, vp, d_mask, d_old, counts, stream);
  if (rc) return rc;
  const size_t blocks = (n_words + kDiffFusedThreads - 1) / kDiffFusedThreads;
  hipLaunchKernelGGL(diff_write_fused_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kDiffFusedThreads), 0,
                     as_stream(stream), reinterpret_cast<unsigned long long*>(d_old),
                     reinterpret_cast<const unsigned long long*>(d_mask), n_words, n_objects,
                     reinterpret_cast<const uint4*>(counts), d_changed_idx, update_old, d_count);
  return check_launch("diff_write_fused_kernel");
}

}  // extern "C"

// batch_types.hpp -- device-visible descriptor layouts shared by the batched
// kernels (batch.hip) and the host code that fills them (host/Serving.cpp).
#pragma once
#include <cstdint>

namespace dpc {

// One (frustum, group) pair of a batch: the frustum's index in the batch's
// matrix array and the first 64-bit mask word of its result.
struct KPair {
  uint32_t frustum;
  uint32_t pad;
  uint64_t mask_word;
};
static_assert(sizeof(KPair) == 16, "KPair layout");

// One request of a serving batch: the whole scene against pairs
// [pair_begin, pair_begin + pair_count).  cost_begin = sum of tiles * pair_count
// over the earlier entries (the kernel splits the total cost evenly over waves).
struct ServeEntry {
  uint32_t scene;
  uint32_t pair_begin;
  uint32_t pair_count;
  uint32_t tiles;  // ceil(objects / (64 * objects_per_lane))
  uint64_t cost_begin;
  uint64_t objects;
};
static_assert(sizeof(ServeEntry) == 32, "ServeEntry layout");

// Launches one serving batch; every pointer is device memory.
int launch_serve(const void* d_entries, uint32_t n_entries, uint64_t total_cost, const void* d_pairs,
                 const void* d_scenes, const float* d_vps, uint64_t* d_masks, uint32_t objects_per_lane,
                 uint32_t waves_per_cu, void* stream);

// Scatters count delta matrices (16 floats each, row-major) into row planes,
// reading the indices and matrices from pinned host memory (device-visible
// pointers) with system-coherent loads; no copy is needed before it.  At most
// kMaxHostScatter matrices into at most kMaxHostScatterLd slots (32-bit byte
// offsets).
constexpr uint32_t kMaxHostScatter = 1u << 25;
constexpr uint32_t kMaxHostScatterLd = (1u << 26) - 1;
int launch_scatter_host(float* d_planes, size_t ld, const uint32_t* idx, const float* src, size_t count,
                        void* stream);

// Scatters count delta matrices that crossed PCIe packed (device memory, 16-byte
// aligned): matrix k as three 16-byte rows in d_packed, (row i's x y z, row 3's
// component i) for i = 0, 1, 2, and its last column (m[3], m[7], m[11], m[15])
// as 16 bytes in d_tails, or, with d_tails null, 0 0 0 1 for every matrix.
// Every value is moved as its bits.
int launch_scatter_affine(float* d_planes, size_t ld, const uint32_t* d_idx, const void* d_packed,
                          const void* d_tails, size_t count, void* stream);

// One request's masks to its caller's buffer in pinned host memory: `words`
// 64-bit words from byte offset src_offset of the batch arena to dst (the
// buffer's device-visible address).
struct CopyOut {
  uint64_t src_offset;
  uint64_t words;
  uint64_t* dst;
  uint64_t pad;
};
static_assert(sizeof(CopyOut) == 32, "CopyOut layout");

// Copies n CopyOut ranges (d_copies: device memory) out of the arena in one
// launch; the device writes the callers' pinned buffers itself, in 512-byte
// wave-wide stores, instead of a DMA transfer per request.
int launch_copy_out(const void* d_copies, uint32_t n, const char* d_arena, uint64_t max_words, void* stream);

}  // namespace dpc

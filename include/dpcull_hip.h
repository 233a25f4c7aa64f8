/*
 * dpcull_hip.h -- C ABI of the MI355X (gfx950) backend for nvpro-pipeline's
 * scene-update path: world-matrix propagation in the transform tree and
 * bounding-box frustum culling.
 *
 * The reference exposes this path as C++ virtual interfaces:
 *   dp::culling::Manager          (dp/culling/Manager.h:108-143)
 *   dp::transform::Tree           (dp/transform/Tree.h:42-129)
 * Those interfaces are mirrored in C++ by include/dpcull/Manager.hpp and
 * include/dpcull/Tree.hpp, which call only the functions below.  Each entry
 * point names the reference code it replaces.
 *
 * Conventions
 *   - Every function returns DPC_OK (0) or a DPC_ERR_* code; nothing throws
 *     across this boundary.  dpc_last_error() returns the calling thread's
 *     last message.
 *   - "d_" pointers are device memory (dpc_malloc or any HIP allocation on the
 *     current device); other pointers are host memory borrowed for the call.
 *   - Matrices are 4x4 FP32, row-major, row-vector convention (v' = v * M), as
 *     dp::math::Mat44f (dp/math/Matmnt.h:199).  On the device they live in
 *     "row planes": plane r (r = 0..3) holds row r of every matrix as float4,
 *     plane r starts at d_planes + r * ld * 4 floats.  ld >= matrix count.
 *   - Object boxes live in two device arrays: box_a[i] = (lower.x, lower.y,
 *     lower.z, extent.x) as float4 and box_b[i] = (extent.y, extent.z) as
 *     float2, where extent = upper - lower (ManagerBitSet.cpp:99-109).
 *   - A visibility mask of n objects is ceil(n/64) 64-bit words, little-endian,
 *     bit i = object i (dp::util::BitArray layout read as uint32 words,
 *     dp/culling/cpu/src/ManagerImpl.cpp:477,517).  Bits >= n are 0.  Mask
 *     pointers must be 8-byte aligned.
 *   - stream arguments are hipStream_t values passed as void* (NULL = the
 *     null stream).  Launch functions are asynchronous on that stream.
 *   - FP32 arithmetic follows the reference's evaluation order with no fused
 *     multiply-add, so masks are bit-identical to dp::culling::cpu and world
 *     matrices to dp::transform::Tree::compute.
 */
#ifndef DPCULL_HIP_H
#define DPCULL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPC_OK 0
#define DPC_ERR_INVALID_ARGUMENT 1
#define DPC_ERR_HIP 2
#define DPC_ERR_OUT_OF_MEMORY 3
#define DPC_ERR_OVER_BUDGET 4
#define DPC_ERR_NOT_FOUND 5
#define DPC_ERR_STATE 6
#define DPC_ERR_INTERNAL 7 /* a launch schedule failed its own consistency check (a library bug) */

#define DPC_ABI_VERSION 2

/* ---- runtime plumbing (no reference counterpart: the reference's CPU path
 *      needs none, its GL sibling uses dp/gl buffers, opengl/src/GroupImpl.cpp:68-183) */
int dpc_abi_version(void);
const char* dpc_last_error(void);
/* Test hook (no reference counterpart): after `after_calls` more checked HIP
 * calls of this library succeed, the next one reports hip_error (a hipError_t
 * value, as DPC_ERR_HIP) instead of its own result; after_calls < 0 turns the
 * injection off.  Lets tests drive every error path of the host code (the C++
 * mirrors turn it into std::runtime_error) and check the state afterwards. */
int dpc_debug_inject_error(int64_t after_calls, int hip_error);
/* Calibration probe (no reference counterpart): the device's issue rate of
 * independent FP32 adds (op 0) or multiplies (op 1), unfused and unpacked as
 * every operation of the culling kernels is.  Eight independent accumulators per
 * lane for `iters` rounds, 32 waves per CU, five launches timed with HIP events
 * on `stream`; *lane_ops_per_s = lane operations per second.  The batched
 * cull's issue fraction is read against it. */
int dpc_probe_valu_rate(int op, uint32_t iters, double* lane_ops_per_s, void* stream);
int dpc_device_info(int device, char* name, size_t name_len, int* cu_count, int* clock_khz,
                    size_t* total_mem, int* l2_bytes);
int dpc_set_device(int device);
int dpc_malloc(void** d_ptr, size_t bytes);
int dpc_free(void* d_ptr);
int dpc_host_alloc(void** h_ptr, size_t bytes); /* pinned host memory */
int dpc_host_free(void* h_ptr);
int dpc_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream);
int dpc_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream);
int dpc_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* Fills n 32-bit words with value using a kernel (safe inside graph capture). */
int dpc_fill_u32(uint32_t* d_dst, uint32_t value, size_t n, void* stream);
int dpc_stream_create(void** stream);
int dpc_stream_destroy(void* stream);
int dpc_stream_sync(void* stream);
int dpc_device_sync(void);
int dpc_event_create(void** event);
int dpc_event_destroy(void* event);
int dpc_event_record(void* event, void* stream);
int dpc_event_sync(void* event);
int dpc_event_elapsed_ms(void* start, void* stop, float* ms);
/* hipGraph capture of work issued on a stream, and replay. */
int dpc_graph_begin(void* stream);
int dpc_graph_end(void* stream, void** graph_exec);
int dpc_graph_launch(void* graph_exec, void* stream);
int dpc_graph_destroy(void* graph_exec);

/* ---- layout conversion ------------------------------------------------------
 * Reference matrices are an array with a byte stride
 * (GroupBitSet::setMatrices, dp/culling/src/GroupBitSet.cpp:121-135;
 * dp::transform::Tree::getWorldMatrices, dp/transform/Tree.h:81).
 * d_aos holds count matrices of 64 bytes at stride_bytes apart (a copy of
 * that array on the device); they are written to row planes d_planes[0..count). */
int dpc_matrices_from_aos(const void* d_aos, size_t stride_bytes, size_t count, float* d_planes, size_t ld,
                          void* stream);
/* Inverse: row planes -> AoS, optionally through an index map
 * (d_aos[k] = plane matrix d_slot[k]; d_slot NULL = identity). */
int dpc_matrices_to_aos(const float* d_planes, size_t ld, const uint32_t* d_slot, size_t count, void* d_aos,
                        size_t stride_bytes, void* stream);
/* Scatter of dirty matrices ("delta"): plane matrix d_idx[k] = d_src[k] (16 floats
 * each).  Device-side form of Tree::updateLocalMatrix (dp/transform/Tree.h:85)
 * and of rewriting a matrix then calling Manager::groupMatrixChanged
 * (dp/culling/Manager.h:125). */
int dpc_matrices_scatter(float* d_planes, size_t ld, const uint32_t* d_idx, const float* d_src, size_t count,
                         void* stream);
/* d_out[k] = d_table[d_idx[k]], or 0xFFFFFFFF when d_idx[k] >= table_len.  Maps
 * transform indices to tree slots for groups bound to a GPU tree. */
int dpc_gather_u32(const uint32_t* d_table, size_t table_len, const uint32_t* d_idx, size_t n, uint32_t* d_out,
                   void* stream);
/* ManagerBitSet::objectSetBoundingBox (dp/culling/src/ManagerBitSet.cpp:99-109)
 * for n boxes at once: d_lower_upper is n x 6 floats (lower.xyz, upper.xyz);
 * extent = upper - lower in FP32. */
int dpc_boxes_from_lower_upper(const float* d_lower_upper, size_t n, float* d_box_a, float* d_box_b,
                               void* stream);

/* ---- frustum culling --------------------------------------------------------
 * dp::culling::cpu::ManagerImpl::cull (dp/culling/cpu/src/ManagerImpl.cpp:467-518)
 * with GroupCPU::updateOBBs (:114-162) fused in: for each object i < n,
 *   M     = matrix d_xform[i] (or matrix i when d_xform is NULL)
 *   bit i = isVisible(vp, OBB(box_i, M))            (:199-229, :263-289)
 * vp is host memory (16 floats), passed by value to the kernel. */
int dpc_cull(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
             const float* d_planes, size_t ld, const float vp[16], uint64_t* d_mask, void* stream);

/* Launch shape of dpc_cull (process-wide; chosen by the autotuner).
 * waves_per_cu = 0 (the default): one wave per 64-object tile over a full grid;
 * waves_per_cu > 0: a persistent grid of that many waves per CU, each wave
 * streaming over tiles with the next tile's loads in flight. */
typedef struct dpc_cull_config {
  uint32_t waves_per_cu;
} dpc_cull_config;
int dpc_cull_get_config(dpc_cull_config* cfg);
int dpc_cull_set_config(const dpc_cull_config* cfg);

/* ---- batched multi-frustum culling -------------------------------------------
 * Many cull(group, result, viewProjection) calls (dp/culling/Manager.h:139) in
 * one persistent launch.  Objects live in "scenes" (device SoA arrays as above);
 * a group is a contiguous object range of one scene; a pair asks for
 * (frustum, group) and names where its mask starts in d_masks (in 64-bit
 * words).  Each object's matrix and box are read once per launch and tested
 * against every frustum paired with its group.  Groups of different scenes can
 * share one launch (the serving loop batches requests across resident scenes). */
typedef struct dpc_scene_desc {
  const float* box_a;     /* n_objects x float4 */
  const float* box_b;     /* n_objects x float2 */
  const uint32_t* xform;  /* n_objects matrix indices, or NULL: object i uses matrix i */
  const float* planes;    /* 4 row planes of ld matrices */
  uint64_t ld;
  uint64_t n_objects;
} dpc_scene_desc;

typedef struct dpc_group_range {
  uint32_t scene;
  uint32_t reserved;
  uint64_t first_object;
  uint64_t count;
} dpc_group_range;

typedef struct dpc_cull_pair {
  uint32_t frustum;
  uint32_t group;
  uint64_t mask_word_offset;
} dpc_cull_pair;

typedef struct dpc_batch_config {
  uint32_t objects_per_lane; /* 1, 2 or 4: a wave owns 64*objects_per_lane objects */
  uint32_t block_threads;    /* 64, 128, 256 */
  uint32_t waves_per_cu;     /* persistent grid size = CUs * waves_per_cu / waves per block */
  uint32_t max_frusta_per_item; /* split a tile's frustum list into chunks of this size (0 = no split) */
} dpc_batch_config;

typedef struct dpc_batch_plan dpc_batch_plan;

/* Builds the work list (tiles sorted by cost) and uploads it.  cfg NULL = defaults. */
int dpc_batch_plan_create(const dpc_group_range* groups, size_t n_groups, const dpc_cull_pair* pairs,
                          size_t n_pairs, uint32_t n_frusta, const dpc_batch_config* cfg, dpc_batch_plan** plan);
int dpc_batch_plan_destroy(dpc_batch_plan* plan);
/* Total mask words the pairs address (max offset + words), tests = sum of |group| over pairs. */
int dpc_batch_plan_info(const dpc_batch_plan* plan, uint64_t* mask_words, uint64_t* tests, uint64_t* work_items);
/* Binds the scene table (host memory, copied): checks every paired group lies
 * inside its scene, then uploads the table.  Rebind when arrays move. */
int dpc_batch_plan_bind(dpc_batch_plan* plan, const dpc_scene_desc* scenes, size_t n_scenes);
/* d_vps: n_frusta x 16 floats in device memory (update it in place between graph
 * replays); d_masks holds mask_words 64-bit words, at least the plan's mask_words. */
int dpc_cull_batch(const dpc_batch_plan* plan, const float* d_vps, uint64_t* d_masks, size_t mask_words,
                   void* stream);
/* The per-view form of the same work: one launch per frustum over its groups (h_vps in host memory). */
int dpc_cull_per_view(const dpc_batch_plan* plan, const float* h_vps, uint64_t* d_masks, size_t mask_words,
                      void* stream);

/* ---- transform tree propagation ----------------------------------------------
 * dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166) on a tree
 * laid out level-major: the entries of level l occupy slots
 * [h_level_offsets[l], h_level_offsets[l+1]) (any order inside a level whose
 * entries do not read what other entries of it write),
 * and d_parent_slot[s] is the slot of entry s's parent.  For each level in
 * order, for each slot s of the level:
 *     if dirty_world[parent] || dirty_local[s]:  W[s] = L[s] * W[parent]; dirty_world[s] = 1
 * d_dirty_local NULL means every local matrix is dirty (the first compute after
 * construction).  d_dirty_world (one byte per slot) must be 0 on entry for
 * every slot the levels write; the caller clears it after consuming the event
 * (Tree.cpp:162-165).  Entries of one level must not write a slot another entry
 * of the same level reads (Tree's compile step splits such levels). */
int dpc_tree_propagate(const float* d_local, float* d_world, size_t ld, const uint32_t* d_parent_slot,
                       const uint64_t* h_level_offsets, size_t n_levels, const uint8_t* d_dirty_local,
                       uint8_t* d_dirty_world, void* stream);
/* The same propagation through a launch schedule built once per layout.
 * dpc_tree_plan_create reads the layout on the host (h_parent_slot: ld entries;
 * the level offsets as above) and keeps a device copy of the parent slots.
 * Consecutive passes in which every entry's parent lies in the previous pass
 * and which are ordered by parent slot become tile runs: one launch in which a
 * workgroup takes a slot range of the run's last pass together with every
 * ancestor of it in the run's earlier passes, computes them pass by pass with
 * each pass's results in LDS for the next, and writes the slots it owns
 * (neighbouring tiles share at most one slot per pass and compute it both).
 * Where no such run exists, small passes share a single-workgroup launch and
 * any other pass is one launch.  tile_slots is the workgroup size of the tile
 * runs (slots per step).  Every tile run is checked against the properties
 * its kernel relies on before anything is uploaded (DPC_ERR_INTERNAL if one
 * fails).  dpc_tree_propagate_plan has dpc_tree_propagate's semantics for
 * that layout.  n_levels = 0 (a tree holding only its root) gives a plan that
 * launches nothing. */
typedef struct dpc_tree_plan dpc_tree_plan;
int dpc_tree_plan_create(const uint32_t* h_parent_slot, size_t ld, const uint64_t* h_level_offsets, size_t n_levels,
                         uint32_t tile_slots, dpc_tree_plan** plan);  /* tile_slots: 256, 512, 1024 or 0 (1024) */
int dpc_tree_plan_destroy(dpc_tree_plan* plan);
/* launches per propagation, how many are tile runs, and their total tiles */
int dpc_tree_plan_info(const dpc_tree_plan* plan, uint32_t* launches, uint32_t* tile_runs, uint64_t* tiles);
int dpc_tree_propagate_plan(const dpc_tree_plan* plan, const float* d_local, float* d_world,
                            const uint8_t* d_dirty_local, uint8_t* d_dirty_world, void* stream);

/* Packs one byte per slot into BitArray words through a slot->index map:
 * bit d_index_of_slot[s] of d_bits = d_bytes[s] != 0 (d_bits must be cleared first;
 * it holds n_bits bits, and indices >= n_bits are skipped).
 * Produces the EventWorldMatricesChanged bitset (dp/transform/Tree.h:48-60). */
int dpc_bytes_to_bits_indexed(const uint8_t* d_bytes, const uint32_t* d_index_of_slot, size_t n_slots,
                              uint32_t* d_bits, size_t n_bits, void* stream);

/* World-matrix inverse transpose: dp::math::invertTranspose
 * (dp/math/Matmnt.h:1171-1201), which the renderer applies to every changed
 * world matrix (ShaderManagerTransformsRiXFx::updateTransforms,
 * dp/sg/renderer/rix/gl/src/ShaderManagerRiXFx.cpp:162-182).  For each slot
 * s < n with d_dirty == NULL or d_dirty[s] != 0 (the tree's dirty-world bytes):
 * when det != 0 the rows of inverse(W[s])^T go to d_out_planes (row planes,
 * out_ld) and d_ok[s] = 1; when det == 0 d_out_planes is left as it was and
 * d_ok[s] = 0 (the reference returns false).  d_ok may be NULL. */
int dpc_inverse_transpose(const float* d_planes, size_t ld, const uint8_t* d_dirty, size_t n, float* d_out_planes,
                          size_t out_ld, uint8_t* d_ok, void* stream);

/* ---- result diff (ResultBitSet::updateChanged, dp/culling/src/ResultBitSet.cpp:61-108) ---
 * changed = old ^ new over n bits; the indices of changed bits are written to
 * d_changed_idx in ascending order (BitArray::traverseBits order) and their
 * count to *d_count; then old = new if update_old != 0.  d_workspace needs
 * dpc_result_diff_workspace(n) bytes. */
size_t dpc_result_diff_workspace(size_t n_bits);
int dpc_result_diff(uint64_t* d_old, const uint64_t* d_new, size_t n_bits, uint32_t* d_changed_idx,
                    uint32_t* d_count, int update_old, void* d_workspace, void* stream);

/* dpc_result_diff with no workspace argument and no allocation of any kind.
 * No reference counterpart: it is for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, hosts that forbid
 * hipMalloc).  Its per-block counts live in a fixed 256 KiB device array of the
 * library, one per device, held until the process ends; n_bits above
 * DPC_RESULT_DIFF_NOALLOC_MAX_BITS (2^30) is rejected with
 * DPC_ERR_INVALID_ARGUMENT.  Every call on one device shares that array, so
 * calls that overlap on one device (concurrent streams, or threads with their
 * own streams) race: serialise them, e.g. on one stream.  Results are those of
 * dpc_result_diff. */
#define DPC_RESULT_DIFF_NOALLOC_MAX_BITS 1073741824ull
int dpc_result_diff_noalloc(uint64_t* d_old, const uint64_t* d_new, size_t n_bits, uint32_t* d_changed_idx,
                            uint32_t* d_count, int update_old, void* stream);

/* dpc_cull followed by dpc_result_diff of its mask against d_old: the tail of
 * cpu::ManagerImpl::cull, updateChanged(visible) (dp/culling/cpu/src/ManagerImpl.cpp:517,
 * dp/culling/src/ResultBitSet.cpp:61-108).  d_mask receives the visibility bits
 * as from dpc_cull; d_changed_idx, *d_count and (with update_old) d_old are
 * then as from dpc_result_diff(d_old, d_mask, n_objects, ...).  The cull also
 * counts each word's changed bits, so up to 2^22 objects this takes two
 * launches instead of three (above that, the result diff's own launches
 * follow the cull).  d_old and d_mask must not overlap; d_workspace needs
 * dpc_cull_diff_workspace(n_objects) bytes, 16-byte aligned. */
size_t dpc_cull_diff_workspace(size_t n_objects);
int dpc_cull_diff(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
                  const float* d_planes, size_t ld, const float vp[16], uint64_t* d_mask, uint64_t* d_old,
                  uint32_t* d_changed_idx, uint32_t* d_count, int update_old, void* d_workspace, void* stream);

/* ---- group bounding box (ManagerBitSet::calculateBoundingBox, dp/culling/src/ManagerBitSet.cpp:268-306)
 * World-space AABB of every object's eight transformed corners; d_out6 =
 * lower.xyz, upper.xyz as the reference's Box3f(min, max) returns them.
 * d_workspace needs dpc_bbox_workspace(n) bytes, 8-byte aligned. */
size_t dpc_bbox_workspace(size_t n_objects);
int dpc_bbox(const float* d_box_a, const float* d_box_b, const uint32_t* d_xform, size_t n_objects,
             const float* d_planes, size_t ld, float* d_out6, void* d_workspace, void* stream);

/* ---- serving loop --------------------------------------------------------------
 * No reference counterpart: the reference culls one scene per frame on the
 * render thread (dp/sg/renderer/rix/gl/src/SceneRendererImpl.cpp:333-356).  A
 * server keeps scenes resident on the device and runs a scheduler thread that
 * batches cull requests of different scenes into one launch (continuous
 * batching: requests that arrive while batches run join the next free stream).
 * A request = apply a matrix delta to its scene, then cull the whole scene
 * against 1..64 frusta.  Requests of one scene run in submission order, one per
 * batch, on one stream while any of them is in flight; requests are admitted
 * while the device memory they need fits the budget left after the resident
 * scenes (a quarter of the budget, at most 1 GiB, is the in-flight arena,
 * split over streams x batches_per_stream batches).  Each batch also has a
 * pinned host arena of the same size, at most 32 MiB, where its descriptors and
 * copied deltas are gathered for one transfer. */
typedef struct dpc_server dpc_server;

typedef struct dpc_server_config {
  uint32_t max_scenes;          /* resident scenes (<= 64) */
  uint32_t max_batch_requests;  /* requests per launch (<= 64) */
  uint32_t streams;             /* ordering domains (HIP streams), 1..8 */
  uint32_t objects_per_lane;    /* kernel shape: 1 or 2 */
  uint32_t waves_per_cu;        /* kernel shape: persistent waves per CU */
  uint32_t batches_per_stream;  /* batches in flight per stream (0: 2, at most 4) */
  uint32_t direct_delta_reads;  /* 1: a delta whose two arrays each lie in one pinned host
                                   allocation is read by the device directly, with no
                                   copy; 0: every delta is copied to the device first */
  uint64_t hbm_budget_bytes;    /* resident scenes + in-flight requests */
} dpc_server_config;

typedef struct dpc_request_info {
  double submit_us;   /* all times: microseconds on the server's steady clock */
  double admit_us;
  double done_us;
  uint32_t batch;     /* batch sequence number */
  uint32_t batch_requests;
  uint64_t mask_words; /* n_frusta x ceil(n_objects / 64) */
  int32_t status;     /* DPC_OK or an error code */
  uint32_t reserved;
} dpc_request_info;

typedef struct dpc_server_stats {
  uint64_t requests_done;
  uint64_t batches;
  uint64_t tests;            /* object-frustum tests */
  uint64_t delta_matrices;
  uint64_t scene_bytes;      /* device bytes of resident scenes */
  uint64_t peak_inflight_bytes;
  double busy_us;            /* sum over batches of launch-to-completion time */
  double issue_us;           /* scheduler-thread time spent issuing batches (copies, launches) */
  double issue_max_us;       /* the longest single batch issue */
  /* the longest time one batch spent in each part of its issue: [0] delta
   * copies, [1] delta scatter launches, [2] descriptor copy, [3] cull launch,
   * [4] mask copies and completion event */
  double issue_part_max_us[5];
  uint64_t delta_direct;     /* requests whose delta the device read straight from pinned host memory */
  uint64_t delta_packed;     /* copied delta matrices scattered from their first three columns alone
                                (their request's every matrix had last column 0 0 0 1 exactly) */
} dpc_server_stats;

int dpc_server_create(const dpc_server_config* cfg, dpc_server** server);
int dpc_server_destroy(dpc_server* server);
/* Uploads a scene (host arrays, copied): lower_upper n x 6, matrices n_matrices x 16
 * (row-major), xform n indices or NULL (object i uses matrix i).
 * DPC_ERR_OVER_BUDGET when the scene does not fit. */
int dpc_server_add_scene(dpc_server* server, const float* lower_upper, uint64_t n_objects, const float* matrices,
                         uint64_t n_matrices, const uint32_t* xform, uint32_t* scene_id);
int dpc_server_remove_scene(dpc_server* server, uint32_t scene_id);
/* Queues a request.  vps (n_frusta x 16), delta_idx (n_delta distinct matrix
 * indices) and delta_mats (n_delta x 16) are read while the request is being
 * launched; masks_out (at least n_frusta x ceil(n_objects / 64) words,
 * frustum-major, the dp::util::BitArray layout) receives the result.  All four
 * are borrowed until dpc_server_wait returns for the request.  With
 * direct_delta_reads, delta arrays that each lie within one pinned host
 * allocation (dpc_host_alloc) are read by the device directly, with no copy;
 * others are copied first: gathered by the host with the batch's other deltas
 * into the batch's pinned arena and transferred in one DMA copy (matrix arrays
 * above 32 MiB, or batches whose deltas exceed the arena, are transferred
 * from the caller's arrays instead).  Gathered matrices travel as their first
 * three columns; their last columns (m[3], m[7], m[11], m[15]) travel too only
 * in a batch where some gathered matrix's last column differs, bit for bit,
 * from 0 0 0 1; either way the scene receives every matrix bit for bit.  A masks_out
 * whose mask_capacity_words lie within one pinned host allocation is written by
 * the device directly (one copy-out launch per batch, no DMA transfer); any
 * other masks_out receives a copy. */
int dpc_server_submit(dpc_server* server, uint32_t scene_id, const float* vps, uint32_t n_frusta,
                      const uint32_t* delta_idx, const float* delta_mats, uint32_t n_delta, uint64_t* masks_out,
                      uint64_t mask_capacity_words, uint64_t* request_id);
/* Waits for a request (timeout_ms < 0: no limit) and fills info.  The request is
 * forgotten afterwards.  Returns the request's status; DPC_ERR_STATE on timeout.
 * A request whose batch failed to launch (status != DPC_OK) applied no delta. */
int dpc_server_wait(dpc_server* server, uint64_t request_id, double timeout_ms, dpc_request_info* info);
int dpc_server_stats_get(dpc_server* server, dpc_server_stats* stats);
/* Copies a scene's current matrices (every delta of its completed requests
 * applied) to the host: n_matrices (the scene's count) x 16 floats, row-major,
 * as dpc_server_add_scene takes them.  DPC_ERR_STATE while the scene has queued
 * or in-flight requests.  The server's scheduler waits for the copy. */
int dpc_server_read_matrices(dpc_server* server, uint32_t scene_id, float* matrices, uint64_t n_matrices);
/* The server's clock (microseconds), to schedule arrivals against. */
double dpc_server_now_us(dpc_server* server);

#ifdef __cplusplus
}
#endif

#endif /* DPCULL_HIP_H */

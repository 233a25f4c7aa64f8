// Serving.cpp -- a request loop over resident scenes (dpc_server_* in
// include/dpcull_hip.h).
//
// Scenes are uploaded once and stay in device memory.  Callers submit requests
// (scene, frusta, matrix delta) from any thread; a scheduler thread forms
// batches continuously.  Each stream holds up to batches_per_stream batches in
// flight, each with an arena of its own.  Whenever an arena is free, the
// scheduler takes for its stream, in arrival order, at most one request per
// scene, until the arena is full.  A scene with requests in flight is bound to
// the stream that runs them and joins only batches of that stream: stream order
// then keeps each scene's requests in submission order (each sees the deltas
// of the earlier ones) without waiting for them to finish.  A batch copies the
// deltas to the device (gathered on the host into one copy, each matrix as its
// first three columns when the batch's last columns are all 0 0 0 1) and
// scatters them into the scenes' matrices, tests every
// scene of the batch against its request's frusta in one launch
// (dpc::launch_serve), then copies each request's masks to its caller's
// buffer.  Batches on different streams overlap on the device.
//
// The HIP calls of a batch are made by its stream's issuer thread, one per
// stream, in the order the scheduler formed the stream's batches: a call that
// blocks (the runtime's copy path can, for milliseconds) holds up that stream
// only, never the scheduler or the other streams.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <immintrin.h>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../batch_types.hpp"
#include "dpcull_hip.h"
#include "trace.hpp"

namespace dpc {
int set_error(int code, const char* fmt, ...);  // runtime.hip: the thread's dpc_last_error() text
}

namespace {

constexpr uint32_t kMaxScenes = 64;
constexpr uint32_t kMaxBatch = 64;
constexpr uint32_t kMaxFrusta = 64;
// the pinned staging area of one batch's descriptors (entries, pairs, frusta, copy-outs)
constexpr uint64_t kHostArena = 4096 + kMaxBatch * 32 + kMaxBatch * kMaxFrusta * (16 + 64) + kMaxBatch * 32 + 4 * 256;
// ... which also stages the batch's deltas when they fit: the batch then
// crosses PCIe in one copy instead of two per request
constexpr uint64_t kHostStageMax = 32ull << 20;
// a delta's matrices up to this size are staged; a larger array is copied on its own
constexpr uint64_t kStageMatsMax = kHostStageMax;

uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// One host copy of a gather: bytes from src to dst or, with tail set, bytes / 64
// matrices packed by pack_affine.
struct Piece {
  char* dst;
  const char* src;
  uint64_t bytes;
  char* tail = nullptr;
  std::atomic<bool>* nonaffine = nullptr;  // set when a packed matrix's last column is not 0 0 0 1
};

// Packs matrices [j0, j1) of a delta (16 floats each, row-major) as three
// 16-byte rows in packed, (row i's x y z, row 3's component i) for i = 0, 1, 2,
// and the last column (m[3], m[7], m[11], m[15]) as 16 bytes in tail; packed
// and tail are 16-byte aligned.  Every value is moved as its bits (shuffles
// only), with streaming stores: the DMA copy reads the arena from memory, and
// the stores need not read its lines first.  Returns whether some matrix's
// last column differs, bit for bit, from 0 0 0 1 (0.0f, 0.0f, 0.0f, 1.0f).
bool pack_affine(const char* src, char* packed, char* tail, uint64_t j0, uint64_t j1)
{
  const __m128i affine = _mm_set_epi32(0x3f800000, 0, 0, 0);
  __m128i other = _mm_setzero_si128();
  for (uint64_t j = j0; j < j1; ++j) {
    const char* m = src + 64 * j;
    const __m128 r0 = _mm_castsi128_ps(_mm_loadu_si128(reinterpret_cast<const __m128i*>(m)));
    const __m128 r1 = _mm_castsi128_ps(_mm_loadu_si128(reinterpret_cast<const __m128i*>(m + 16)));
    const __m128 r2 = _mm_castsi128_ps(_mm_loadu_si128(reinterpret_cast<const __m128i*>(m + 32)));
    const __m128 r3 = _mm_castsi128_ps(_mm_loadu_si128(reinterpret_cast<const __m128i*>(m + 48)));
    // row i's z and row 3's component i, then (x, y, z, row 3's component i)
    const __m128 z0 = _mm_shuffle_ps(r0, r3, _MM_SHUFFLE(0, 0, 2, 2));
    const __m128 z1 = _mm_shuffle_ps(r1, r3, _MM_SHUFFLE(1, 1, 2, 2));
    const __m128 z2 = _mm_shuffle_ps(r2, r3, _MM_SHUFFLE(2, 2, 2, 2));
    const __m128 p0 = _mm_shuffle_ps(r0, z0, _MM_SHUFFLE(2, 0, 1, 0));
    const __m128 p1 = _mm_shuffle_ps(r1, z1, _MM_SHUFFLE(2, 0, 1, 0));
    const __m128 p2 = _mm_shuffle_ps(r2, z2, _MM_SHUFFLE(2, 0, 1, 0));
    // the w of each row: (r0.w, r1.w, r2.w, r3.w)
    const __m128 u = _mm_unpackhi_ps(r0, r1);  // r0.z r1.z r0.w r1.w
    const __m128 v = _mm_unpackhi_ps(r2, r3);  // r2.z r3.z r2.w r3.w
    const __m128i w = _mm_castps_si128(_mm_movehl_ps(v, u));
    __m128i* dp = reinterpret_cast<__m128i*>(packed + 48 * j);
    _mm_stream_si128(dp, _mm_castps_si128(p0));
    _mm_stream_si128(dp + 1, _mm_castps_si128(p1));
    _mm_stream_si128(dp + 2, _mm_castps_si128(p2));
    _mm_stream_si128(reinterpret_cast<__m128i*>(tail + 16 * j), w);
    other = _mm_or_si128(other, _mm_xor_si128(w, affine));
  }
  _mm_sfence();  // the streaming stores are visible before this thread reports done
  return _mm_movemask_epi8(_mm_cmpeq_epi8(other, _mm_setzero_si128())) != 0xFFFF;
}

// Helper threads shared by the issuers' gathers.  A gather offers every part
// but its first to the helpers, runs the first itself, then takes back and
// runs each part that no helper has started yet, and waits for the parts that
// helpers are running: it never waits for a helper to become free.
class GatherPool {
 public:
  // Starts up to `helpers` threads (a thread that cannot be started is left out).
  void start(unsigned helpers)
  {
    try {
      for (unsigned k = 0; k < helpers; ++k) threads_.emplace_back([this] { loop(); });
    } catch (...) {
    }
  }
  ~GatherPool() { stop(); }
  // Joins the helpers (again: nothing to do); no gather may be running.
  void stop()
  {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
    threads_.clear();
  }
  unsigned helpers() const { return static_cast<unsigned>(threads_.size()); }
  // Runs fn(0), ..., fn(parts - 1), on this thread and the helpers.
  void run(unsigned parts, const std::function<void(unsigned)>& fn)
  {
    if (parts <= 1 || threads_.empty()) {
      for (unsigned t = 0; t < parts; ++t) fn(t);
      return;
    }
    Job job{&fn, parts - 1};
    {
      std::lock_guard<std::mutex> g(mu_);
      for (unsigned t = 1; t < parts; ++t) tasks_.push_back(Task{&job, t});
    }
    cv_work_.notify_all();
    fn(0);
    for (;;) {
      unsigned part = 0;
      {
        std::lock_guard<std::mutex> g(mu_);
        const auto it = std::find_if(tasks_.begin(), tasks_.end(), [&](const Task& k) { return k.job == &job; });
        if (it == tasks_.end()) break;
        part = it->part;
        tasks_.erase(it);
      }
      fn(part);
      std::lock_guard<std::mutex> g(mu_);
      --job.pending;
    }
    // (a helper touches the job for the last time when it counts its part done, under the lock)
    std::unique_lock<std::mutex> lock(mu_);
    cv_done_.wait(lock, [&] { return job.pending == 0; });
  }

 private:
  struct Job {
    const std::function<void(unsigned)>* fn;
    unsigned pending;  // parts not yet done, the first one excepted
  };
  struct Task {
    Job* job;
    unsigned part;
  };
  void loop()
  {
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
      cv_work_.wait(lock, [&] { return stop_ || !tasks_.empty(); });
      if (tasks_.empty()) return;  // (stopping)
      const Task k = tasks_.front();
      tasks_.pop_front();
      lock.unlock();
      (*k.job->fn)(k.part);
      lock.lock();
      if (--k.job->pending == 0) cv_done_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Task> tasks_;
  std::vector<std::thread> threads_;
  bool stop_ = false;
};

constexpr unsigned kGatherHelpers = 4;

// Copies (or packs) every piece.  From 512 KiB on, the source bytes are split
// into equal runs of at least 256 KiB for this thread and the pool's helpers:
// one thread's memcpy would hold up the batch's transfer.  A packed piece is
// split at whole matrices: a run packs the matrices that start within it.
void gather(GatherPool& pool, const std::vector<Piece>& pieces)
{
  uint64_t total = 0;
  for (const Piece& p : pieces) total += p.bytes;
  constexpr uint64_t kSplitMin = 512ull << 10, kPerPartMin = 256ull << 10;
  const unsigned parts =
      total < kSplitMin ? 1u : static_cast<unsigned>(std::min<uint64_t>(1 + pool.helpers(), total / kPerPartMin));
  // part t copies the bytes [t * total / parts, (t + 1) * total / parts) of the concatenated pieces
  pool.run(parts, [&](unsigned t) {
    const uint64_t lo = total * t / parts, hi = total * (t + 1) / parts;
    uint64_t at = 0;
    for (const Piece& p : pieces) {
      const uint64_t a = std::max(lo, at), b = std::min(hi, at + p.bytes);
      if (a < b && !p.tail) {
        std::memcpy(p.dst + (a - at), p.src + (a - at), b - a);
      } else if (a < b && pack_affine(p.src, p.dst, p.tail, (a - at + 63) / 64, (b - at + 63) / 64)) {
        p.nonaffine->store(true, std::memory_order_relaxed);
      }
      at += p.bytes;
      if (at >= hi) break;
    }
  });
}

struct Scene {
  bool live = false;
  bool ready = false;     // uploaded: may be scheduled
  uint32_t inflight = 0;  // requests in batches that have not completed
  int stream = -1;        // the stream those batches run on (-1: none)
  uint64_t n = 0, n_mat = 0, bytes = 0;
  float* box_a = nullptr;
  float* box_b = nullptr;
  float* planes = nullptr;
  uint32_t* xform = nullptr;
  uint32_t pending = 0;  // queued requests
};

struct Request {
  uint64_t id = 0;
  uint32_t scene = 0;
  const float* vps = nullptr;
  uint32_t n_frusta = 0;
  const uint32_t* didx = nullptr;
  const float* dmats = nullptr;
  uint32_t n_delta = 0;
  // both delta arrays lie in pinned host memory: the device reads them through
  // these pointers, with no copy
  const uint32_t* didx_dev = nullptr;
  const float* dmats_dev = nullptr;
  uint64_t* masks_out = nullptr;
  uint64_t* masks_dev = nullptr;  // masks_out lies in pinned host memory: the device writes it, at this address
  uint64_t words = 0;       // mask words per frustum
  uint64_t dev_bytes = 0;   // arena bytes this request needs
  bool delta_packed = false;  // its delta crossed PCIe as 48 bytes per matrix
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint32_t batch = 0, batch_requests = 0;
  int status = DPC_OK;
  bool done = false;
};

// One batch in flight: its arena, completion event and requests.
struct StreamSlot {
  hipStream_t stream = nullptr;  // borrowed from dpc_server::streams
  int stream_index = 0;
  hipEvent_t done = nullptr;
  char* d_arena = nullptr;
  char* h_arena = nullptr;  // pinned, h_cap bytes: the device arena's first h_cap bytes staged on the host
  uint64_t cap = 0;
  uint64_t h_cap = 0;
  bool busy = false;     // formed and not yet completed
  bool issued = false;   // its issuer has recorded the completion event
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<Scene> snapshot;  // the batch's scenes at forming (arrays stable while requests are in flight)
  double t_launch = 0;
  uint64_t bytes = 0;
  uint64_t tests = 0;
};

}  // namespace

struct dpc_server;

namespace {

// The thread that makes the HIP calls of one stream's batches, in order.
struct Issuer {
  std::thread thread;
  std::condition_variable cv;
  std::deque<size_t> queue;  // slots to issue
};

}  // namespace

struct dpc_server {
  dpc_server_config cfg{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv_done;
  std::condition_variable cv_work;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  std::vector<Scene> scenes;
  dpc_scene_desc* d_scene_table = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<StreamSlot> slots;
  uint64_t next_id = 1;
  uint32_t batch_seq = 0;
  uint64_t scene_bytes = 0;
  uint64_t arena_total = 0;
  bool stop = false;
  std::thread worker;
  GatherPool pool;  // helpers of the issuers' gathers
  std::vector<std::unique_ptr<Issuer>> issuers;  // one per stream
  bool issuers_stop = false;
  dpc_server_stats stats{};
  int device = 0;

  double now() const
  {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  }
  void run();
  void issue_loop(int stream_index);
  bool form_batch(size_t k);
  void issue_batch(StreamSlot& slot);
  void complete(StreamSlot& slot);
};

namespace {

int fail(int code, const std::string& msg) { return dpc::set_error(code, "%s", msg.c_str()); }

// The device-visible address of [p, p + bytes) when the whole range lies in
// one pinned host allocation (hipHostMalloc), else nullptr.
const void* pinned_range(const void* p, size_t bytes)
{
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeHost || !a.devicePointer) {
    (void)hipGetLastError();
    return nullptr;
  }
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, reinterpret_cast<hipDeviceptr_t>(const_cast<void*>(p))) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
  if (q < b || bytes > size || q - b > size - bytes) return nullptr;
  return a.devicePointer;
}

int hip_fail(hipError_t e, const char* what)
{
  return fail(DPC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// Forms a batch on slot k and hands it to the slot's issuer.  Called by the
// scheduler with the lock held.
bool dpc_server::form_batch(size_t k)
{
  StreamSlot& slot = slots[k];
  std::vector<std::shared_ptr<Request>> pick;
  std::vector<bool> scene_taken(scenes.size(), false);
  uint64_t bytes = 0;
  for (auto it = pending.begin(); it != pending.end() && pick.size() < cfg.max_batch_requests;) {
    const auto& r = *it;
    Scene& sc = scenes[r->scene];
    const bool other_stream = sc.inflight != 0 && sc.stream != slot.stream_index;
    if (!sc.ready || other_stream || scene_taken[r->scene]) {
      scene_taken[r->scene] = true;  // later requests of this scene wait behind this one
      ++it;
      continue;
    }
    if (bytes + r->dev_bytes > slot.cap) break;  // admission: the arena (memory budget) is full
    scene_taken[r->scene] = true;
    bytes += r->dev_bytes;
    pick.push_back(r);
    it = pending.erase(it);
  }
  if (pick.empty()) return false;
  const uint32_t batch = ++batch_seq;
  const double t_admit = now();
  slot.snapshot.clear();
  for (auto& r : pick) {
    scenes[r->scene].inflight++;
    scenes[r->scene].stream = slot.stream_index;
    scenes[r->scene].pending--;
    r->t_admit = t_admit;
    r->batch = batch;
    r->batch_requests = static_cast<uint32_t>(pick.size());
    slot.snapshot.push_back(scenes[r->scene]);
  }
  slot.reqs = std::move(pick);
  slot.busy = true;
  slot.issued = false;
  slot.t_launch = t_admit;
  Issuer& is = *issuers[slot.stream_index];
  is.queue.push_back(k);
  is.cv.notify_one();
  return true;
}

// Makes the HIP calls of the batch on `slot` (its issuer thread, without the
// lock: the slot's batch belongs to the issuer until it is marked issued).
void dpc_server::issue_batch(StreamSlot& slot)
{
  dpcull::TraceRange trace("serve batch issue");
  const std::vector<std::shared_ptr<Request>>& pick = slot.reqs;
  const std::vector<Scene>& snapshot = slot.snapshot;
  const double t_issue0 = now();

  // Arena layout: the batch descriptors (entries, pairs, frusta; staged in the
  // pinned host arena at the same offsets), then per request the delta indices
  // and matrices, then the masks.
  const uint32_t K = cfg.objects_per_lane;
  hipStream_t s = slot.stream;
  hipError_t e = hipSuccess;
  uint32_t n_pairs = 0;
  for (auto& r : pick) n_pairs += r->n_frusta;
  uint64_t off = 0;
  const uint64_t entries_off = off;
  off = align_up(off + sizeof(dpc::ServeEntry) * pick.size(), 256);
  const uint64_t pairs_off = off;
  off = align_up(off + sizeof(dpc::KPair) * n_pairs, 256);
  const uint64_t vps_off = off;
  off = align_up(off + 64ull * n_pairs, 256);
  const uint64_t copies_off = off;
  off = align_up(off + sizeof(dpc::CopyOut) * pick.size(), 256);
  const uint64_t desc_end = off;
  // The copied deltas: every index array and the matrices of deltas up to
  // kStageMatsMax bytes first (the part staged on the host, below), then the
  // larger matrix arrays, each copied on its own.  When the staged part fits
  // the host arena its matrices are packed: their first three columns (48
  // bytes) up to packed_end, then their last columns (16 bytes), which cross
  // PCIe only in a batch where some matrix's last column is not 0 0 0 1.
  std::vector<uint64_t> idx_off(pick.size()), mat_off(pick.size()), tail_off(pick.size());
  uint64_t packed_end = 0, stage_end = 0;
  auto lay_out = [&](bool pack) {
    uint64_t o = desc_end;
    for (size_t k = 0; k < pick.size(); ++k) {
      const Request& r = *pick[k];
      const uint64_t n = r.didx_dev ? 0 : r.n_delta;
      idx_off[k] = o;
      o = align_up(o + 4ull * n, 256);
      if (64ull * n <= kStageMatsMax) {
        mat_off[k] = o;
        o = align_up(o + (pack ? 48ull : 64ull) * n, 256);
      }
    }
    packed_end = o;
    for (size_t k = 0; k < pick.size() && pack; ++k) {
      const Request& r = *pick[k];
      const uint64_t n = r.didx_dev ? 0 : r.n_delta;
      if (64ull * n <= kStageMatsMax) {
        tail_off[k] = o;
        o = align_up(o + 16ull * n, 256);
      }
    }
    stage_end = o;
    for (size_t k = 0; k < pick.size(); ++k) {
      const Request& r = *pick[k];
      const uint64_t n = r.didx_dev ? 0 : r.n_delta;
      if (64ull * n > kStageMatsMax) {
        mat_off[k] = o;
        o = align_up(o + 64ull * n, 256);
      }
    }
    return o;
  };
  off = lay_out(true);
  const bool staged = stage_end <= slot.h_cap;
  if (!staged) off = lay_out(false);
  const uint64_t masks_off = off;
  auto* entries = reinterpret_cast<dpc::ServeEntry*>(slot.h_arena + entries_off);
  auto* pairs = reinterpret_cast<dpc::KPair*>(slot.h_arena + pairs_off);
  auto* vps = reinterpret_cast<float*>(slot.h_arena + vps_off);
  auto* copies = reinterpret_cast<dpc::CopyOut*>(slot.h_arena + copies_off);
  uint32_t n_copies = 0;
  uint64_t copy_words = 0;
  std::vector<uint64_t> mask_off(pick.size());  // byte offset of each request's masks in the device arena
  uint64_t cost = 0, mask_word = 0, tests = 0;
  uint32_t p = 0;
  for (size_t k = 0; k < pick.size(); ++k) {
    const Request& r = *pick[k];
    const Scene& sc = snapshot[k];
    dpc::ServeEntry& en = entries[k];
    en.scene = r.scene;
    en.pair_begin = p;
    en.pair_count = r.n_frusta;
    en.tiles = static_cast<uint32_t>((sc.n + 64ull * K - 1) / (64ull * K));
    en.cost_begin = cost;
    en.objects = sc.n;
    cost += static_cast<uint64_t>(en.tiles) * r.n_frusta;
    mask_off[k] = masks_off + 8 * mask_word;
    for (uint32_t f = 0; f < r.n_frusta; ++f, ++p) {
      pairs[p].frustum = p;
      pairs[p].pad = 0;
      pairs[p].mask_word = mask_word;
      mask_word += r.words;
      std::memcpy(vps + 16ull * p, r.vps + 16ull * f, 64);
    }
    tests += sc.n * r.n_frusta;
    if (r.masks_dev) {
      copies[n_copies++] = dpc::CopyOut{mask_off[k], r.words * r.n_frusta, r.masks_dev, 0};
      copy_words = std::max<uint64_t>(copy_words, r.words * r.n_frusta);
    }
  }
  const uint64_t end = masks_off + 8 * mask_word;
  int rc = DPC_OK;
  if (end > slot.cap) rc = fail(DPC_ERR_OVER_BUDGET, "batch larger than its arena");
  double t_part[6];
  double copy_us = 0.0;
  t_part[0] = now();
  // The descriptors, the index arrays and the smaller matrix arrays (packed)
  // staged together in the pinned host arena when they fit: one copy for them,
  // then the larger matrix arrays on their own, then the scatters.  Otherwise
  // each delta is copied on its own, before its scatter.
  std::vector<std::atomic<bool>> nonaffine(pick.size());  // (all false)
  if (staged) {
    std::vector<Piece> pieces;
    for (size_t k = 0; k < pick.size(); ++k) {
      const Request& r = *pick[k];
      if (!r.n_delta || r.didx_dev) continue;
      pieces.push_back(Piece{slot.h_arena + idx_off[k], reinterpret_cast<const char*>(r.didx), 4ull * r.n_delta});
      if (64ull * r.n_delta <= kStageMatsMax)
        pieces.push_back(Piece{slot.h_arena + mat_off[k], reinterpret_cast<const char*>(r.dmats), 64ull * r.n_delta,
                               slot.h_arena + tail_off[k], &nonaffine[k]});
    }
    gather(pool, pieces);
    bool tails = false;
    for (auto& f : nonaffine) tails = tails || f.load(std::memory_order_relaxed);
    if (rc == DPC_OK) {
      e = hipMemcpyAsync(slot.d_arena, slot.h_arena, tails ? stage_end : packed_end, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) rc = hip_fail(e, "batch upload");
    }
    copy_us = now() - t_part[0];
  }
  // deltas: copy to the device (unless staged above) and scatter into the scenes' matrices
  for (size_t k = 0; k < pick.size() && rc == DPC_OK; ++k) {
    const Request& r = *pick[k];
    if (!r.n_delta) continue;
    if (r.didx_dev) {  // pinned: the scatter reads the caller's arrays directly
      rc = dpc::launch_scatter_host(snapshot[k].planes, snapshot[k].n_mat, r.didx_dev, r.dmats_dev, r.n_delta, s);
      continue;
    }
    if (staged && 64ull * r.n_delta <= kStageMatsMax) {
      const bool tail = nonaffine[k].load(std::memory_order_relaxed);
      rc = dpc::launch_scatter_affine(snapshot[k].planes, snapshot[k].n_mat,
                                      reinterpret_cast<const uint32_t*>(slot.d_arena + idx_off[k]),
                                      slot.d_arena + mat_off[k], tail ? slot.d_arena + tail_off[k] : nullptr,
                                      r.n_delta, s);
      if (!tail) pick[k]->delta_packed = true;
      continue;
    }
    {
      const double tc = now();
      if (!staged) e = hipMemcpyAsync(slot.d_arena + idx_off[k], r.didx, 4ull * r.n_delta, hipMemcpyHostToDevice, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(slot.d_arena + mat_off[k], r.dmats, 64ull * r.n_delta, hipMemcpyHostToDevice, s);
      copy_us += now() - tc;
      if (e != hipSuccess) {
        rc = hip_fail(e, "delta upload");
        break;
      }
    }
    rc = dpc_matrices_scatter(snapshot[k].planes, snapshot[k].n_mat,
                              reinterpret_cast<const uint32_t*>(slot.d_arena + idx_off[k]),
                              reinterpret_cast<const float*>(slot.d_arena + mat_off[k]), r.n_delta, s);
  }
  t_part[2] = now();
  t_part[1] = t_part[0] + copy_us;  // parts 0 and 1: copies, then the scatter launches (interleaved in time)
  if (rc == DPC_OK && !staged) {
    e = hipMemcpyAsync(slot.d_arena, slot.h_arena, desc_end, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "descriptor upload");
  }
  t_part[3] = now();
  if (rc == DPC_OK) {
    rc = dpc::launch_serve(slot.d_arena + entries_off, static_cast<uint32_t>(pick.size()), cost,
                           slot.d_arena + pairs_off, d_scene_table, reinterpret_cast<const float*>(slot.d_arena + vps_off),
                           reinterpret_cast<uint64_t*>(slot.d_arena + masks_off), K, cfg.waves_per_cu, s);
  }
  t_part[4] = now();
  // each request's masks go straight to the caller's buffer: written by the
  // device in one launch for every buffer in pinned memory, copied otherwise
  if (rc == DPC_OK) rc = dpc::launch_copy_out(slot.d_arena + copies_off, n_copies, slot.d_arena, copy_words, s);
  for (size_t k = 0; k < pick.size() && rc == DPC_OK; ++k) {
    const Request& r = *pick[k];
    if (r.masks_dev) continue;
    e = hipMemcpyAsync(r.masks_out, slot.d_arena + mask_off[k], 8 * r.words * r.n_frusta, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = hip_fail(e, "mask download");
  }
  if (rc == DPC_OK) {
    e = hipEventRecord(slot.done, s);
    if (e != hipSuccess) rc = hip_fail(e, "completion event");
  }
  if (rc != DPC_OK) (void)hipStreamSynchronize(s);  // nothing may still read the arena when the batch ends
  const double t_issue1 = now();
  t_part[5] = t_issue1;
  std::lock_guard<std::mutex> lock(mu);
  stats.issue_us += t_issue1 - t_issue0;
  stats.issue_max_us = std::max(stats.issue_max_us, t_issue1 - t_issue0);
  for (int q = 0; q < 5; ++q) stats.issue_part_max_us[q] = std::max(stats.issue_part_max_us[q], t_part[q + 1] - t_part[q]);
  slot.bytes = end;
  slot.tests = tests;
  stats.peak_inflight_bytes = std::max<uint64_t>(stats.peak_inflight_bytes, end);
  if (rc != DPC_OK) {
    for (auto& r : slot.reqs) r->status = rc;
    complete(slot);
  } else {
    slot.issued = true;
  }
  cv_work.notify_all();  // the scheduler polls issued batches
}

void dpc_server::issue_loop(int stream_index)
{
  (void)hipSetDevice(device);
  Issuer& is = *issuers[stream_index];
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    is.cv.wait(lock, [&] { return issuers_stop || !is.queue.empty(); });
    if (is.queue.empty()) break;  // (stopping, and nothing left to issue)
    const size_t k = is.queue.front();
    is.queue.pop_front();
    lock.unlock();
    issue_batch(slots[k]);
    lock.lock();
  }
}

// Called with the lock held once the slot's completion event has fired.
void dpc_server::complete(StreamSlot& slot)
{
  const double t = now();
  for (size_t k = 0; k < slot.reqs.size(); ++k) {
    Request& r = *slot.reqs[k];
    if (r.status == DPC_OK) {
      stats.delta_matrices += r.n_delta;
      if (r.n_delta && r.didx_dev) stats.delta_direct++;
      if (r.delta_packed) stats.delta_packed += r.n_delta;
    }
    r.t_done = t;
    r.done = true;
    Scene& sc = scenes[r.scene];
    if (--sc.inflight == 0) sc.stream = -1;
    stats.requests_done++;
  }
  stats.batches++;
  stats.tests += slot.tests;
  stats.busy_us += t - slot.t_launch;
  slot.reqs.clear();
  slot.snapshot.clear();
  slot.busy = false;
  slot.issued = false;
  cv_done.notify_all();
}

void dpc_server::run()
{
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    bool progressed = false;
    bool any_busy = false;
    for (auto& slot : slots) {
      if (slot.busy && !slot.issued) {
        any_busy = true;  // its issuer is making the calls
      } else if (slot.busy) {
        const hipError_t q = hipEventQuery(slot.done);
        if (q == hipSuccess) {
          complete(slot);
          progressed = true;
        } else if (q != hipErrorNotReady) {
          // the batch cannot be observed to finish (the device reported an
          // error): its requests fail instead of waiting for ever
          (void)hipGetLastError();
          for (auto& r : slot.reqs) r->status = DPC_ERR_HIP;
          complete(slot);
          progressed = true;
        } else {
          any_busy = true;
        }
      }
    }
    if (!pending.empty()) {
      // free arenas of the least busy streams first: a request of a scene with
      // nothing in flight then starts on an idle stream instead of queueing
      // behind another batch (a scene bound to a busy stream can only go there)
      std::vector<int> load(streams.size(), 0);
      for (const auto& slot : slots) load[slot.stream_index] += slot.busy ? 1 : 0;
      std::vector<bool> tried(slots.size(), false);
      while (!pending.empty()) {
        size_t best = slots.size();
        for (size_t k = 0; k < slots.size(); ++k)
          if (!slots[k].busy && !tried[k] &&
              (best == slots.size() || load[slots[k].stream_index] < load[slots[best].stream_index]))
            best = k;
        if (best == slots.size()) break;
        tried[best] = true;
        if (form_batch(best)) {
          load[slots[best].stream_index]++;
          progressed = true;
          any_busy = true;
        }
      }
    }
    if (stop && pending.empty() && !any_busy) break;
    if (progressed) continue;
    if (any_busy) {
      // a batch is running: poll its event without sleeping (latency matters)
      lock.unlock();
      std::this_thread::yield();
      lock.lock();
    } else {
      // (a deadline on the system clock: libstdc++ waits for it with
      // pthread_cond_timedwait, which ThreadSanitizer follows; a clock step
      // only moves this re-check, since new work notifies the scheduler)
      cv_work.wait_until(lock, std::chrono::system_clock::now() + std::chrono::microseconds(200));
    }
  }
}

extern "C" {

int dpc_server_create(const dpc_server_config* cfg, dpc_server** out)
{
  if (!cfg || !out) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_create: null argument");
  *out = nullptr;
  if (cfg->max_scenes == 0 || cfg->max_scenes > kMaxScenes || cfg->max_batch_requests == 0 ||
      cfg->max_batch_requests > kMaxBatch || cfg->streams == 0 || cfg->streams > 8 || cfg->batches_per_stream > 4 ||
      cfg->direct_delta_reads > 1 ||
      (cfg->objects_per_lane != 1 && cfg->objects_per_lane != 2) || cfg->hbm_budget_bytes < (64ull << 20))
    return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_create: bad configuration");
  auto srv = std::make_unique<dpc_server>();
  srv->cfg = *cfg;
  (void)hipGetDevice(&srv->device);
  srv->scenes.resize(cfg->max_scenes);
  // a quarter of the budget (at most 1 GiB) is the in-flight request arena, split
  // over the streams; each stream has a device arena and a pinned host twin
  srv->arena_total = std::min<uint64_t>(cfg->hbm_budget_bytes / 4, 1ull << 30);
  const uint32_t depth = cfg->batches_per_stream ? cfg->batches_per_stream : 2;
  srv->cfg.batches_per_stream = depth;
  const uint64_t per = align_up(srv->arena_total / (cfg->streams * depth), 4096);
  hipError_t e = hipMalloc(&srv->d_scene_table, sizeof(dpc_scene_desc) * cfg->max_scenes);
  if (e != hipSuccess) return hip_fail(e, "scene table");
  std::vector<dpc_scene_desc> zero(cfg->max_scenes, dpc_scene_desc{});
  e = hipMemcpy(srv->d_scene_table, zero.data(), sizeof(dpc_scene_desc) * cfg->max_scenes, hipMemcpyHostToDevice);
  srv->streams.assign(cfg->streams, nullptr);
  for (auto& st : srv->streams)
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  srv->slots.resize(static_cast<size_t>(cfg->streams) * depth);
  for (size_t k = 0; k < srv->slots.size(); ++k) {
    StreamSlot& sl = srv->slots[k];
    sl.stream_index = static_cast<int>(k % cfg->streams);
    sl.stream = srv->streams[sl.stream_index];
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sl.d_arena), per);
    sl.h_cap = std::max(kHostArena, std::min(per, kHostStageMax));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_arena), sl.h_cap, hipHostMallocDefault);
    sl.cap = per;
  }
  if (e != hipSuccess) {
    int rc = hip_fail(e, "dpc_server_create");
    dpc_server_destroy(srv.release());
    return rc;
  }
  dpc_server* s = srv.release();
  for (uint32_t k = 0; k < cfg->streams; ++k) s->issuers.push_back(std::make_unique<Issuer>());
  for (uint32_t k = 0; k < cfg->streams; ++k) {
    const int index = static_cast<int>(k);
    s->issuers[k]->thread = std::thread([s, index] { s->issue_loop(index); });
  }
  s->pool.start(kGatherHelpers);
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return DPC_OK;
}

int dpc_server_destroy(dpc_server* s)
{
  if (!s) return DPC_OK;
  if (s->worker.joinable()) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
  }
  // the scheduler has ended with every batch completed: the issuers have nothing left
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->issuers_stop = true;
  }
  for (auto& is : s->issuers) {
    is->cv.notify_all();
    if (is->thread.joinable()) is->thread.join();
  }
  s->pool.stop();  // (no issuer is left to gather)
  for (auto& st : s->streams)
    if (st) (void)hipStreamSynchronize(st);
  for (auto& sl : s->slots) {
    if (sl.d_arena) (void)hipFree(sl.d_arena);
    if (sl.h_arena) (void)hipHostFree(sl.h_arena);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  for (auto& st : s->streams)
    if (st) (void)hipStreamDestroy(st);
  for (auto& sc : s->scenes) {
    if (!sc.live) continue;
    (void)hipFree(sc.box_a);
    (void)hipFree(sc.box_b);
    (void)hipFree(sc.planes);
    if (sc.xform) (void)hipFree(sc.xform);
  }
  if (s->d_scene_table) (void)hipFree(s->d_scene_table);
  delete s;
  return DPC_OK;
}

int dpc_server_add_scene(dpc_server* s, const float* lower_upper, uint64_t n, const float* matrices, uint64_t n_mat,
                         const uint32_t* xform, uint32_t* scene_id)
{
  if (!s || !scene_id || (n && !lower_upper) || (n_mat && !matrices) || n == 0 || n_mat == 0)
    return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_add_scene: bad argument");
  if (!xform && n_mat < n) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_add_scene: n_matrices < n_objects");
  const uint64_t bytes = n * 24 + n_mat * 64 + (xform ? n * 4 : 0);
  uint32_t id;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->scene_bytes + bytes + s->arena_total > s->cfg.hbm_budget_bytes)
      return fail(DPC_ERR_OVER_BUDGET, "dpc_server_add_scene: the scene does not fit the budget");
    id = 0;
    while (id < s->scenes.size() && s->scenes[id].live) ++id;
    if (id == s->scenes.size()) return fail(DPC_ERR_OVER_BUDGET, "dpc_server_add_scene: all scene slots are in use");
    s->scenes[id].live = true;    // reserve the slot
    s->scenes[id].ready = false;  // not schedulable until uploaded
    s->scene_bytes += bytes;
  }
  Scene sc;
  sc.n = n;
  sc.n_mat = n_mat;
  sc.bytes = bytes;
  sc.live = true;
  sc.ready = true;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&sc.box_a), n * 16);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc.box_b), n * 8);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc.planes), n_mat * 64);
  if (e == hipSuccess && xform) e = hipMalloc(reinterpret_cast<void**>(&sc.xform), n * 4);
  void *d_lu = nullptr, *d_aos = nullptr;
  if (e == hipSuccess) e = hipMalloc(&d_lu, n * 24);
  if (e == hipSuccess) e = hipMalloc(&d_aos, n_mat * 64);
  if (e == hipSuccess) e = hipMemcpy(d_lu, lower_upper, n * 24, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_aos, matrices, n_mat * 64, hipMemcpyHostToDevice);
  if (e == hipSuccess && xform) e = hipMemcpy(sc.xform, xform, n * 4, hipMemcpyHostToDevice);
  int rc = e == hipSuccess ? DPC_OK : hip_fail(e, "dpc_server_add_scene: upload");
  if (rc == DPC_OK) rc = dpc_boxes_from_lower_upper(static_cast<const float*>(d_lu), n, sc.box_a, sc.box_b, nullptr);
  if (rc == DPC_OK) rc = dpc_matrices_from_aos(d_aos, 64, n_mat, sc.planes, n_mat, nullptr);
  dpc_scene_desc desc{sc.box_a, sc.box_b, sc.xform, sc.planes, n_mat, n};
  if (rc == DPC_OK) {
    e = hipMemcpy(s->d_scene_table + id, &desc, sizeof desc, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "scene table");
  }
  if (rc == DPC_OK) {
    e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "dpc_server_add_scene");
  }
  if (d_lu) (void)hipFree(d_lu);
  if (d_aos) (void)hipFree(d_aos);
  std::lock_guard<std::mutex> g(s->mu);
  if (rc != DPC_OK) {
    if (sc.box_a) (void)hipFree(sc.box_a);
    if (sc.box_b) (void)hipFree(sc.box_b);
    if (sc.planes) (void)hipFree(sc.planes);
    if (sc.xform) (void)hipFree(sc.xform);
    s->scenes[id] = Scene{};
    s->scene_bytes -= bytes;
    return rc;
  }
  s->scenes[id] = sc;
  s->stats.scene_bytes = s->scene_bytes;
  *scene_id = id;
  return DPC_OK;
}

int dpc_server_remove_scene(dpc_server* s, uint32_t id)
{
  if (!s) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_remove_scene: null server");
  Scene sc;
  {
    std::unique_lock<std::mutex> lock(s->mu);
    if (id >= s->scenes.size() || !s->scenes[id].live) return fail(DPC_ERR_NOT_FOUND, "no such scene");
    if (!s->scenes[id].ready || s->scenes[id].inflight || s->scenes[id].pending)
      return fail(DPC_ERR_STATE, "dpc_server_remove_scene: the scene has requests in flight");
    sc = s->scenes[id];
    s->scenes[id] = Scene{};
    s->scene_bytes -= sc.bytes;
    s->stats.scene_bytes = s->scene_bytes;
  }
  dpc_scene_desc zero{};
  (void)hipMemcpy(s->d_scene_table + id, &zero, sizeof zero, hipMemcpyHostToDevice);
  (void)hipFree(sc.box_a);
  (void)hipFree(sc.box_b);
  (void)hipFree(sc.planes);
  if (sc.xform) (void)hipFree(sc.xform);
  return DPC_OK;
}

int dpc_server_submit(dpc_server* s, uint32_t scene_id, const float* vps, uint32_t n_frusta, const uint32_t* didx,
                      const float* dmats, uint32_t n_delta, uint64_t* masks_out, uint64_t mask_capacity_words,
                      uint64_t* request_id)
{
  if (!s || !request_id || !vps || !masks_out || n_frusta == 0 || n_frusta > kMaxFrusta ||
      (n_delta && (!didx || !dmats)))
    return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_submit: bad argument");
  auto r = std::make_shared<Request>();
  r->scene = scene_id;
  r->vps = vps;
  r->n_frusta = n_frusta;
  r->didx = didx;
  r->dmats = dmats;
  r->n_delta = n_delta;
  r->masks_out = masks_out;
  if (mask_capacity_words && mask_capacity_words <= (~0ull >> 4))
    r->masks_dev = static_cast<uint64_t*>(const_cast<void*>(pinned_range(masks_out, 8 * mask_capacity_words)));
  if (s->cfg.direct_delta_reads && n_delta && n_delta <= dpc::kMaxHostScatter) {
    const void* di = pinned_range(didx, 4ull * n_delta);
    const void* dm = di ? pinned_range(dmats, 64ull * n_delta) : nullptr;
    if (di && dm) {
      r->didx_dev = static_cast<const uint32_t*>(di);
      r->dmats_dev = static_cast<const float*>(dm);
    }
  }
  std::lock_guard<std::mutex> g(s->mu);
  // (a scene still being uploaded has no id for its caller yet, and no sizes)
  if (scene_id >= s->scenes.size() || !s->scenes[scene_id].live || !s->scenes[scene_id].ready)
    return fail(DPC_ERR_NOT_FOUND, "no such scene");
  const Scene& sc = s->scenes[scene_id];
  if (sc.n_mat > dpc::kMaxHostScatterLd) {  // the direct scatter addresses the planes with 32-bit offsets
    r->didx_dev = nullptr;
    r->dmats_dev = nullptr;
  }
  // delta indices are checked here, on the host, so the scatter never writes outside the scene
  for (uint32_t k = 0; k < n_delta; ++k)
    if (didx[k] >= sc.n_mat) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_submit: delta index out of range");
  r->words = (sc.n + 63) / 64;
  if (mask_capacity_words < r->words * n_frusta)
    return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_submit: the mask buffer is smaller than n_frusta x ceil(n/64) words");
  const uint64_t staged_delta = r->didx_dev ? 0 : n_delta;  // a pinned delta takes no arena space
  // (the last term: a batch's alignment to 256 bytes of its four descriptor
  // regions and of this request's three delta regions: indices, and matrices
  // packed in two or whole in one)
  r->dev_bytes = align_up(4ull * staged_delta, 256) + align_up(64ull * staged_delta, 256) + sizeof(dpc::ServeEntry) +
                 sizeof(dpc::CopyOut) + (sizeof(dpc::KPair) + 64) * n_frusta + 8 * r->words * n_frusta + 7 * 256;
  if (r->dev_bytes > s->slots[0].cap)
    return fail(DPC_ERR_OVER_BUDGET, "dpc_server_submit: the request alone exceeds the in-flight budget");
  r->id = s->next_id++;
  r->t_submit = s->now();
  s->scenes[scene_id].pending++;
  s->requests[r->id] = r;
  s->pending.push_back(r);
  *request_id = r->id;
  s->cv_work.notify_one();
  return DPC_OK;
}

int dpc_server_wait(dpc_server* s, uint64_t request_id, double timeout_ms, dpc_request_info* info)
{
  if (!s) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_wait: null server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return fail(DPC_ERR_NOT_FOUND, "dpc_server_wait: unknown request");
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->done; };
  if (timeout_ms < 0)
    s->cv_done.wait(lock, ready);
  else if (!s->cv_done.wait_for(lock, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return fail(DPC_ERR_STATE, "dpc_server_wait: timed out");
  if (info) {
    info->submit_us = r->t_submit;
    info->admit_us = r->t_admit;
    info->done_us = r->t_done;
    info->batch = r->batch;
    info->batch_requests = r->batch_requests;
    info->mask_words = r->words * r->n_frusta;
    info->status = r->status;
  }
  s->requests.erase(it);
  return r->status;
}

int dpc_server_stats_get(dpc_server* s, dpc_server_stats* st)
{
  if (!s || !st) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_stats_get: null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  return DPC_OK;
}

int dpc_server_read_matrices(dpc_server* s, uint32_t id, float* matrices, uint64_t n_matrices)
{
  if (!s || !matrices) return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_read_matrices: null argument");
  // (the lock stays held over the copy: no request of the scene can be admitted meanwhile)
  std::lock_guard<std::mutex> g(s->mu);
  if (id >= s->scenes.size() || !s->scenes[id].live || !s->scenes[id].ready)
    return fail(DPC_ERR_NOT_FOUND, "no such scene");
  const Scene& sc = s->scenes[id];
  if (sc.inflight || sc.pending)
    return fail(DPC_ERR_STATE, "dpc_server_read_matrices: the scene has requests in flight");
  if (n_matrices != sc.n_mat)
    return fail(DPC_ERR_INVALID_ARGUMENT, "dpc_server_read_matrices: the scene holds another number of matrices");
  // the four row planes (row r of matrix k at planes[4 * (r * n + k)]), then row-major per matrix
  std::vector<float> rows(16 * sc.n_mat);
  const hipError_t e = hipMemcpy(rows.data(), sc.planes, 64 * sc.n_mat, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "dpc_server_read_matrices");
  for (uint64_t k = 0; k < sc.n_mat; ++k)
    for (int r = 0; r < 4; ++r) std::memcpy(matrices + 16 * k + 4 * r, rows.data() + 4 * (r * sc.n_mat + k), 16);
  return DPC_OK;
}

double dpc_server_now_us(dpc_server* s) { return s ? s->now() : 0.0; }

}  // extern "C"

// serve_stress.cpp -- the serving loop under concurrent callers, on the GPU.
//
// Several submitter threads (each owning some scenes) submit requests with
// matrix deltas, some waiting at once and some later, while another thread
// polls the statistics.  The scheduler and the per-stream issuers run beside
// them.  Every request must succeed, and its masks must equal those of the
// same requests run one at a time on a fresh server (a scene's requests apply
// their deltas in submission order, so each scene's results are fixed), and
// every scene's matrices after either run must equal, bit for bit, its initial
// matrices with every delta applied in order; some deltas hold matrices whose
// last column is not 0 0 0 1 (a negative zero, a NaN, a projective term, a
// homogeneous scale).  Run for both delta paths (copied and read directly from
// pinned memory).
//
// Built with -DDPCULL_HOST_RUNTIME and ThreadSanitizer or AddressSanitizer
// against the host-only runtime of tests/cpp/host_runtime (bin/tsan_serve_host,
// bin/asan_serve_host, on the CPU), the same threads run the server's scheduler
// and issuers with no device: the runtime moves the matrices but computes no
// masks, and the run checks the statuses, the matrices and that the sanitizer
// finds no data race or bad access among the threads.
#include <atomic>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "dpcull_hip.h"

static std::atomic<int> failures{0};
#define EXPECT(cond, ...)                              \
  do {                                                 \
    if (!(cond)) {                                     \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                        \
      std::printf("\n");                               \
      ++failures;                                      \
    }                                                  \
  } while (0)

struct Scene {
  uint64_t n = 0;
  std::vector<float> lower_upper, matrices;
};

struct Req {
  int scene = 0;
  std::vector<float> vps;
  uint32_t n_frusta = 0;
  uint32_t* idx = nullptr;  // pinned (dpc_host_alloc) or pageable
  float* mats = nullptr;
  uint32_t n_delta = 0;
  bool pinned = false;
  std::vector<uint64_t> masks;
  uint64_t* pinned_masks = nullptr;  // every other request: its masks go to pinned memory (written by the device)
  uint64_t* out() { return pinned_masks ? pinned_masks : masks.data(); }
  void collect()  // the masks as the server delivered them
  {
    if (pinned_masks) std::memcpy(masks.data(), pinned_masks, masks.size() * 8);
  }
};

static void identity_into(float* m)
{
  std::memset(m, 0, 64);
  m[0] = m[5] = m[10] = m[15] = 1.f;
}

// A last column other than 0 0 0 1, of one of four kinds.
static void set_last_column(float* m, uint32_t kind)
{
  const uint32_t nan_bits = 0x7fc01234u;  // a quiet NaN with a payload
  switch (kind % 4) {
    case 0: m[3] = -0.0f; break;
    case 1: std::memcpy(&m[15], &nan_bits, 4); break;
    case 2: m[11] = 0.001f; break;
    default: m[15] = 2.0f; break;
  }
}

static std::vector<Scene> make_scenes(std::mt19937& rng, int count)
{
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::vector<Scene> out(count);
  for (auto& sc : out) {
    sc.n = 4000 + rng() % 40000;
    sc.lower_upper.resize(6 * sc.n);
    sc.matrices.resize(16 * sc.n);
    for (uint64_t i = 0; i < sc.n; ++i) {
      for (int c = 0; c < 3; ++c) {
        const float lo = 100.f * u(rng);
        sc.lower_upper[6 * i + c] = lo;
        sc.lower_upper[6 * i + 3 + c] = lo + 0.5f + 2.f * (u(rng) + 1.f);
      }
      float* m = &sc.matrices[16 * i];
      identity_into(m);
      for (int c = 0; c < 3; ++c) m[12 + c] = 20.f * u(rng);
    }
  }
  return out;
}

static std::vector<Req> make_requests(std::mt19937& rng, const std::vector<Scene>& scenes, int count)
{
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::vector<Req> out(count);
  for (int k = 0; k < count; ++k) {
    Req& r = out[k];
    r.scene = static_cast<int>(rng() % scenes.size());
    const Scene& sc = scenes[r.scene];
    r.n_frusta = 1 + rng() % 6;
    r.vps.resize(16 * r.n_frusta);
    for (uint32_t f = 0; f < r.n_frusta; ++f) {
      float* vp = &r.vps[16 * f];
      identity_into(vp);
      const float s = 0.004f + 0.01f * (u(rng) + 1.f);
      vp[0] = vp[5] = vp[10] = s;
      vp[12] = 0.3f * u(rng);
      vp[13] = 0.3f * u(rng);
    }
    // every tenth request rewrites all of its scene's matrices (up to ~3 MB, which
    // the server gathers on several threads)
    const bool whole = k % 10 == 5;
    r.n_delta = (k % 4 == 3) ? 0 : whole ? static_cast<uint32_t>(sc.n) : static_cast<uint32_t>(1 + rng() % (sc.n / 20));
    r.pinned = (k % 3) != 0;
    if (r.n_delta) {
      void *pi = nullptr, *pm = nullptr;
      if (r.pinned) {
        if (dpc_host_alloc(&pi, 4ull * r.n_delta) != DPC_OK || dpc_host_alloc(&pm, 64ull * r.n_delta) != DPC_OK) {
          std::printf("FAIL: dpc_host_alloc: %s\n", dpc_last_error());
          std::exit(1);
        }
      } else {
        pi = new uint32_t[r.n_delta];
        pm = new float[16ull * r.n_delta];
      }
      r.idx = static_cast<uint32_t*>(pi);
      r.mats = static_cast<float*>(pm);
      // distinct indices: d itself for a whole-scene delta; start + 7d (mod n) for
      // d < n / 20 never meets itself again
      const uint64_t start = rng() % sc.n;
      for (uint32_t d = 0; d < r.n_delta; ++d) {
        r.idx[d] = whole ? d : static_cast<uint32_t>((start + 7ull * d) % sc.n);
        float* m = &r.mats[16ull * d];
        identity_into(m);
        for (int c = 0; c < 3; ++c) m[12 + c] = 20.f * u(rng);
        if (k % 7 == 2 && d % 5 == 0) set_last_column(m, d / 5);
      }
    }
    r.masks.assign(r.n_frusta * ((sc.n + 63) / 64), 0);
    if (k % 2 == 0) {
      void* pm = nullptr;
      if (dpc_host_alloc(&pm, r.masks.size() * 8) != DPC_OK) {
        std::printf("FAIL: dpc_host_alloc: %s\n", dpc_last_error());
        std::exit(1);
      }
      r.pinned_masks = static_cast<uint64_t*>(pm);
    }
  }
  return out;
}

static void free_requests(std::vector<Req>& reqs)
{
  for (auto& r : reqs) {
    if (r.pinned_masks) dpc_host_free(r.pinned_masks);
    if (!r.n_delta) continue;
    if (r.pinned) {
      dpc_host_free(r.idx);
      dpc_host_free(r.mats);
    } else {
      delete[] r.idx;
      delete[] r.mats;
    }
  }
}

static dpc_server* make_server(uint32_t streams, uint32_t direct)
{
  dpc_server_config cfg{};
  cfg.max_scenes = 16;
  cfg.max_batch_requests = 8;
  cfg.streams = streams;
  cfg.objects_per_lane = 1;
  cfg.waves_per_cu = 0;
  cfg.batches_per_stream = 2;
  cfg.direct_delta_reads = direct;
  cfg.hbm_budget_bytes = 2ull << 30;
  dpc_server* s = nullptr;
  if (dpc_server_create(&cfg, &s) != DPC_OK) {
    std::printf("FAIL: dpc_server_create: %s\n", dpc_last_error());
    ++failures;
  }
  return s;
}

static std::vector<uint32_t> add_scenes(dpc_server* s, const std::vector<Scene>& scenes)
{
  std::vector<uint32_t> ids;
  for (const auto& sc : scenes) {
    uint32_t id = 0;
    const int rc = dpc_server_add_scene(s, sc.lower_upper.data(), sc.n, sc.matrices.data(), sc.n, nullptr, &id);
    EXPECT(rc == DPC_OK, "add_scene: %s", dpc_last_error());
    ids.push_back(id);
  }
  return ids;
}

// Each scene's matrices with every request's delta applied in request order.
static std::vector<std::vector<float>> replay(const std::vector<Scene>& scenes, const std::vector<Req>& reqs)
{
  std::vector<std::vector<float>> out;
  for (const auto& sc : scenes) out.push_back(sc.matrices);
  for (const auto& r : reqs)
    for (uint32_t d = 0; d < r.n_delta; ++d) std::memcpy(&out[r.scene][16ull * r.idx[d]], r.mats + 16ull * d, 64);
  return out;
}

// The number of scenes whose matrices on the server differ, in any bit, from want.
static size_t matrices_differ(dpc_server* s, const std::vector<uint32_t>& ids,
                              const std::vector<std::vector<float>>& want)
{
  size_t differ = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    std::vector<float> got(want[i].size());
    const int rc = dpc_server_read_matrices(s, ids[i], got.data(), got.size() / 16);
    EXPECT(rc == DPC_OK, "read_matrices: %s", dpc_last_error());
    differ += std::memcmp(got.data(), want[i].data(), 4 * got.size()) != 0;
  }
  return differ;
}

static void run(uint32_t direct)
{
  std::mt19937 rng(17 + direct);
  const std::vector<Scene> scenes = make_scenes(rng, 6);
  std::vector<Req> reqs = make_requests(rng, scenes, 180);
  const std::vector<std::vector<float>> want_matrices = replay(scenes, reqs);

  // one at a time on a server of one stream: the expected masks
  std::vector<std::vector<uint64_t>> want(reqs.size());
  {
    dpc_server* s = make_server(1, direct);
    const auto ids = add_scenes(s, scenes);
    for (size_t k = 0; k < reqs.size(); ++k) {
      Req& r = reqs[k];
      uint64_t id = 0;
      int rc = dpc_server_submit(s, ids[r.scene], r.vps.data(), r.n_frusta, r.idx, r.mats, r.n_delta, r.out(),
                                 r.masks.size(), &id);
      EXPECT(rc == DPC_OK, "submit: %s", dpc_last_error());
      dpc_request_info info{};
      rc = dpc_server_wait(s, id, -1.0, &info);
      EXPECT(rc == DPC_OK && info.status == DPC_OK, "wait: %d", rc);
      r.collect();
      want[k] = r.masks;
    }
    const size_t bad = matrices_differ(s, ids, want_matrices);
    EXPECT(bad == 0, "one at a time: %zu scenes' matrices differ from the replayed deltas", bad);
    dpc_server_destroy(s);
  }

  // concurrently: three submitters, scene s belongs to submitter s % 3, so each
  // scene's requests keep their order
  dpc_server* s = make_server(4, direct);
  const auto ids = add_scenes(s, scenes);
  for (auto& r : reqs) {
    std::fill(r.masks.begin(), r.masks.end(), 0);
    if (r.pinned_masks) std::memset(r.pinned_masks, 0, r.masks.size() * 8);
  }
  std::atomic<bool> done{false};
  std::thread poller([&] {
    dpc_server_stats st{};
    while (!done.load()) {
      EXPECT(dpc_server_stats_get(s, &st) == DPC_OK, "stats");
      std::this_thread::yield();
    }
  });
  std::vector<std::thread> submitters;
  for (int t = 0; t < 3; ++t) {
    submitters.emplace_back([&, t] {
      std::vector<std::pair<uint64_t, size_t>> later;
      for (size_t k = 0; k < reqs.size(); ++k) {
        Req& r = reqs[k];
        if (r.scene % 3 != t) continue;
        uint64_t id = 0;
        const int rc = dpc_server_submit(s, ids[r.scene], r.vps.data(), r.n_frusta, r.idx, r.mats, r.n_delta,
                                         r.out(), r.masks.size(), &id);
        EXPECT(rc == DPC_OK, "submit: %s", dpc_last_error());
        if (k % 3 == 1) {
          later.emplace_back(id, k);
        } else {
          dpc_request_info info{};
          const int w = dpc_server_wait(s, id, -1.0, &info);
          EXPECT(w == DPC_OK && info.status == DPC_OK, "wait: %d", w);
          r.collect();
        }
      }
      for (auto [id, k] : later) {
        dpc_request_info info{};
        const int w = dpc_server_wait(s, id, -1.0, &info);
        EXPECT(w == DPC_OK && info.status == DPC_OK, "wait later: %d", w);
        reqs[k].collect();
      }
    });
  }
  for (auto& th : submitters) th.join();
  done.store(true);
  poller.join();
  dpc_server_stats st{};
  dpc_server_stats_get(s, &st);
  const size_t bad_matrices = matrices_differ(s, ids, want_matrices);
  EXPECT(bad_matrices == 0, "%zu scenes' matrices differ from the replayed deltas", bad_matrices);
  EXPECT(st.delta_packed > 0, "no delta was scattered packed");
  dpc_server_destroy(s);
  size_t differ = 0;
#ifndef DPCULL_HOST_RUNTIME  // (the host-only runtime computes no masks)
  for (size_t k = 0; k < reqs.size(); ++k) differ += reqs[k].masks != want[k];
#endif
  EXPECT(differ == 0, "%zu of %zu requests differ from the one-at-a-time run", differ, reqs.size());
  EXPECT(st.requests_done == reqs.size(), "requests_done %llu", static_cast<unsigned long long>(st.requests_done));
  std::printf("  %s deltas: %zu requests on 6 scenes from 3 threads, %llu batches, %llu read directly, "
              "%llu matrices packed, %zu differ\n",
              direct ? "direct" : "copied", reqs.size(), static_cast<unsigned long long>(st.batches),
              static_cast<unsigned long long>(st.delta_direct), static_cast<unsigned long long>(st.delta_packed),
              differ);
  free_requests(reqs);
}

int main()
{
  run(0);
  run(1);
  std::printf("serve stress: %d failures\n", failures.load());
  return failures ? 1 : 0;
}

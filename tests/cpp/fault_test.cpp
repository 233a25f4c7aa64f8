// fault_test.cpp -- the host mirrors after a failing HIP call, on the GPU.
//
// dpc_debug_inject_error makes the k-th checked HIP call of the library report
// an error.  For k = 0, 1, 2, ... an operation runs with call k failing; when
// the failure reaches the operation it throws std::runtime_error, and the next
// attempt without injection must give the results of a run that never failed
// (and so must an operation whose failing call was one whose status it ignores):
//   Tree::compute after structural changes (the relayout, the plan, the
//     propagation, the event): world matrices equal the oracle's;
//   Tree::compute with moved transforms only: world matrices, and the event
//     holds exactly the moved transforms and their descendants;
//   Manager::cull with changed matrices and objects (uploads, the cull, the
//     result diff): changed lists and visibility equal the CPU manager's,
//     which ran without failures.
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

#include "cpu_manager.hpp"
#include "dpcull/Manager.hpp"
#include "dpcull/Tree.hpp"
#include "dpcull_hip.h"
#include "oracle.h"

using namespace dpcull;
using namespace dpcull::culling;

static int failures = 0;
#define EXPECT(cond, ...)                              \
  do {                                                 \
    if (!(cond)) {                                     \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                        \
      std::printf("\n");                               \
      ++failures;                                      \
    }                                                  \
  } while (0)

constexpr int kInjected = 2;  // hipErrorOutOfMemory

// Runs op with the k-th checked HIP call failing; true when op threw.
template <typename Op>
static bool fails_at(int64_t k, Op op)
{
  dpc_debug_inject_error(k, kInjected);
  bool threw = false;
  try {
    op();
  } catch (std::runtime_error const& e) {
    threw = true;
    if (k < 3) std::printf("    call %lld failed: %s\n", static_cast<long long>(k), e.what());
  } catch (std::exception const& e) {
    std::printf("    call %lld: unexpected exception: %s\n", static_cast<long long>(k), e.what());
    ++failures;
  }
  dpc_debug_inject_error(-1, 0);
  return threw;
}

// The hook itself on one C-ABI call: the call after `k` good ones fails.
static void hook_counts_calls()
{
  void* d = nullptr;
  dpc_malloc(&d, 4096);
  dpc_debug_inject_error(0, kInjected);
  const int rc0 = dpc_fill_u32(static_cast<uint32_t*>(d), 0u, 1024, nullptr);
  const int rc1 = dpc_fill_u32(static_cast<uint32_t*>(d), 0u, 1024, nullptr);
  dpc_debug_inject_error(1, kInjected);
  const int rc2 = dpc_fill_u32(static_cast<uint32_t*>(d), 0u, 1024, nullptr);
  const int rc3 = dpc_fill_u32(static_cast<uint32_t*>(d), 0u, 1024, nullptr);
  dpc_debug_inject_error(-1, 0);
  dpc_device_sync();
  dpc_free(d);
  EXPECT(rc0 == DPC_ERR_HIP && rc1 == DPC_OK && rc2 == DPC_OK && rc3 == DPC_ERR_HIP,
         "hook: return codes %d %d %d %d", rc0, rc1, rc2, rc3);
}

static Mat44f rand_local(std::mt19937& rng)
{
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  Mat44f m = Mat44f::identity();
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = (r == c ? 1.f : 0.f) + 0.2f * u(rng);
  for (int c = 0; c < 3; ++c) m[3][c] = 5.f * u(rng);
  return m;
}

// Every world matrix from scratch (every local dirty): the oracle's answer for
// any history of successful and failed computes with the same local matrices.
static size_t world_mismatches(const transform::Tree& tree, const std::vector<Mat44f>& local,
                               const std::vector<transform::Index>& nodes)
{
  const size_t cap = tree.getTransformCount();
  std::vector<float> loc(16 * cap, 0.f), world(16 * cap, 0.f);
  std::vector<uint8_t> dl(cap, 1), dw(cap, 0);
  for (size_t i = 0; i < cap && i < local.size(); ++i) std::memcpy(&loc[16 * i], local[i].m, 64);
  std::memcpy(world.data(), Mat44f::identity().m, 64);
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> entries;
  for (const auto& l : tree.layout().levels()) {
    for (const auto& e : l) {
      entries.push_back(e.parent);
      entries.push_back(e.transform);
    }
    off.push_back(entries.size() / 2);
  }
  orc_tree_compute(off.size() - 1, off.data(), entries.data(), loc.data(), world.data(), cap, dl.data(), dw.data(),
                   nullptr);
  const Mat44f* w = tree.getWorldMatrices();
  size_t bad = 0;
  for (transform::Index i : nodes) bad += std::memcmp(w[i].m, &world[16 * i], 64) != 0;
  return bad;
}

static void tree_compute_failures()
{
  std::mt19937 rng(3);
  transform::Tree tree;
  std::vector<Mat44f> local(1 << 17, Mat44f{});
  local[0] = Mat44f::identity();
  std::vector<transform::Index> nodes;
  auto add = [&](int count) {
    for (int k = 0; k < count; ++k) {
      const transform::Index parent = nodes.empty() || rng() % 8 == 0 ? 0 : nodes[rng() % nodes.size()];
      const Mat44f m = rand_local(rng);
      const transform::Index i = tree.addTransform(parent, m);
      local[i] = m;
      nodes.push_back(i);
    }
  };
  add(3000);
  tree.compute();
  int threw = 0, k = 0;
  for (; k < 48; ++k) {
    // a structural change (relayout, new plan) and moved transforms, then a compute failing at call k
    add(40);
    for (int j = 0; j < 60; ++j) {
      const transform::Index i = nodes[rng() % nodes.size()];
      local[i] = rand_local(rng);
      tree.updateLocalMatrix(i, local[i]);
    }
    // (a call whose status the code ignores, such as the free of a buffer that
    // grows, takes the injected error without failing the compute)
    if (fails_at(k, [&] { tree.compute(); })) {
      ++threw;
      tree.compute();  // the retry
    }
    const size_t bad = world_mismatches(tree, local, nodes);
    EXPECT(bad == 0, "tree: compute with call %d failing, then retried: %zu world matrices differ", k, bad);
  }
  EXPECT(threw > 10, "tree: only %d injected failures reached compute", threw);
  std::printf("  tree: %d of %d computes failed (call k failing, k = 0..%d) and recovered\n", threw, k, k - 1);
}

// The same without structural changes (no relayout): a compute that fails
// after its propagation was queued leaves dirty-world bytes behind, and the
// retry's event must still hold exactly the matrices the retry changes: the
// moved transforms and their descendants.
static void tree_event_after_failures()
{
  std::mt19937 rng(5);
  transform::Tree tree;
  std::vector<Mat44f> local(1 << 17, Mat44f{});
  local[0] = Mat44f::identity();
  std::vector<transform::Index> nodes, parent_of(1 << 17, 0);
  for (int k = 0; k < 4000; ++k) {
    const transform::Index parent = nodes.empty() || rng() % 8 == 0 ? 0 : nodes[rng() % nodes.size()];
    const Mat44f m = rand_local(rng);
    const transform::Index i = tree.addTransform(parent, m);
    local[i] = m;
    parent_of[i] = parent;
    nodes.push_back(i);
  }
  tree.compute();
  int threw = 0, k = 0;
  for (; k < 24; ++k) {
    std::vector<uint8_t> moved(1 << 17, 0);
    for (int j = 0; j < 30; ++j) {
      const transform::Index i = nodes[rng() % nodes.size()];
      local[i] = rand_local(rng);
      tree.updateLocalMatrix(i, local[i]);
      moved[i] = 1;
    }
    if (fails_at(k, [&] { tree.compute(); })) {
      ++threw;
      tree.compute();
    }
    // nodes are listed parents first: a node changes when it moved or its parent changed
    std::vector<uint8_t> changed(1 << 17, 0);
    for (transform::Index i : nodes) changed[i] = moved[i] | changed[parent_of[i]];
    const auto& ev = tree.getLastChangedWorldMatrices();
    size_t wrong = 0;
    for (transform::Index i : nodes) {
      const bool bit = (i >> 6) < ev.size() && ((ev[i >> 6] >> (i & 63)) & 1);
      wrong += bit != (changed[i] != 0);
    }
    EXPECT(wrong == 0, "tree event: compute with call %d failing, then retried: %zu indices wrong", k, wrong);
    const size_t bad = world_mismatches(tree, local, nodes);
    EXPECT(bad == 0, "tree event: compute with call %d failing, then retried: %zu world matrices differ", k, bad);
  }
  EXPECT(threw > 5, "tree event: only %d injected failures reached compute", threw);
  std::printf("  tree without relayout: %d of %d computes failed and recovered, events exact\n", threw, k);
}

struct Id : Payload {
  explicit Id(int v) : id(v) {}
  int id;
};

static std::vector<int> ids(Manager& m, std::vector<ObjectSharedPtr> const& v)
{
  std::vector<int> r;
  for (auto& o : v) r.push_back(std::static_pointer_cast<Id>(m.objectGetUserData(o))->id);
  return r;
}

static void manager_cull_failures()
{
  std::mt19937 rng(8);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::unique_ptr<Manager> hipm(hip::Manager::create());
  oracle::CpuManager cpum(true);
  Manager* ms[2] = {hipm.get(), &cpum};
  const int M = 256;
  std::vector<Mat44f> mats(M);
  for (auto& m : mats) m = rand_local(rng);
  GroupSharedPtr g[2];
  ResultSharedPtr r[2];
  std::vector<ObjectSharedPtr> objs[2];
  for (int s = 0; s < 2; ++s) {
    g[s] = ms[s]->groupCreate();
    ms[s]->groupSetMatrices(g[s], mats.data(), M, sizeof(Mat44f));
    r[s] = ms[s]->groupCreateResult(g[s]);
  }
  int next_id = 0;
  auto add_objects = [&](int count) {
    for (int k = 0; k < count; ++k) {
      Box3f b;
      for (int c = 0; c < 3; ++c) {
        b.lower[c] = 20.f * u(rng);
        b.upper[c] = b.lower[c] + 0.5f + (u(rng) + 1.f);
      }
      const size_t t = rng() % M;
      for (int s = 0; s < 2; ++s) {
        auto o = ms[s]->objectCreate(std::make_shared<Id>(next_id));
        ms[s]->objectSetBoundingBox(o, b);
        ms[s]->objectSetTransformIndex(o, t);
        ms[s]->groupAddObject(g[s], o);
        objs[s].push_back(o);
      }
      ++next_id;
    }
  };
  add_objects(5000);
  Mat44f vp = Mat44f::identity();
  vp[0][0] = vp[1][1] = vp[2][2] = 0.05f;
  for (int s = 0; s < 2; ++s) ms[s]->cull(g[s], r[s], vp);
  int threw = 0, k = 0;
  for (; k < 40; ++k) {
    // moved matrices, new objects, a new camera; the HIP cull fails at call k
    for (int j = 0; j < 20; ++j) {
      const size_t i = rng() % M;
      mats[i] = rand_local(rng);
      for (int s = 0; s < 2; ++s) ms[s]->groupMatrixChanged(g[s], i);
    }
    if (k % 3 == 0) add_objects(30);
    vp[3][0] = 0.2f * u(rng);
    vp[3][1] = 0.2f * u(rng);
    const bool failed = fails_at(k, [&] { hipm->cull(g[0], r[0], vp); });
    if (failed) {
      ++threw;
      hipm->cull(g[0], r[0], vp);  // the retry
    }
    cpum.cull(g[1], r[1], vp);
    const auto a = ids(*hipm, hipm->resultGetChanged(r[0]));
    const auto b = ids(cpum, cpum.resultGetChanged(r[1]));
    EXPECT(a == b, "manager: cull failing at call %d: changed lists differ (%zu vs %zu)", k, a.size(), b.size());
    size_t bad = 0;
    for (size_t i = 0; i < objs[0].size(); ++i)
      bad += hipm->resultObjectIsVisible(r[0], objs[0][i]) != cpum.resultObjectIsVisible(r[1], objs[1][i]);
    EXPECT(bad == 0, "manager: cull failing at call %d: %zu visibilities differ", k, bad);
  }
  EXPECT(threw > 3, "manager: only %d injected failures reached cull", threw);
  std::printf("  manager: %d of %d culls failed (call k failing, k = 0..%d) and recovered\n", threw, k, k - 1);
}

int main()
{
  hook_counts_calls();
  tree_compute_failures();
  tree_event_after_failures();
  manager_cull_failures();
  std::printf("fault injection: %d failures\n", failures);
  return failures ? 1 : 0;
}

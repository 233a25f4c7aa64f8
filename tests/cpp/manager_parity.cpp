// manager_parity.cpp -- the HIP culling manager (dpcull::culling::hip::Manager)
// against the CPU restatement of dp::culling::cpu::ManagerImpl (oracle/cpu_manager.hpp).
//
// 1. Hand-traced scenario: expected changed lists and visibility derived by hand
//    from dp/culling/src/{GroupBitSet,ResultBitSet}.cpp and
//    cpu/src/ManagerImpl.cpp (default-visible results, swap-remove remap, the
//    OBB cache that ignores box changes of grouped objects).  Both the oracle and
//    the HIP manager must produce them, in both cache modes.
// 2. Random traces: thousands of adds, removes, box / index / matrix changes,
//    culls with several results and bounding boxes, HIP vs oracle, both modes.
// 3. cullBatch equals the same sequence of cull() calls.
// 4. A group bound to a GPU transform tree equals the oracle fed the tree's
//    world matrices.
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "cpu_manager.hpp"
#include "dpcull/Manager.hpp"
#include "dpcull/Tree.hpp"
#include "self_front.hpp"

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

struct Id : Payload {
  explicit Id(int v) : id(v) {}
  int id;
};

static int id_of(Manager& m, ObjectSharedPtr const& o) { return std::static_pointer_cast<Id>(m.objectGetUserData(o))->id; }

static std::vector<int> ids(Manager& m, std::vector<ObjectSharedPtr> const& v)
{
  std::vector<int> r;
  for (auto& o : v) r.push_back(id_of(m, o));
  return r;
}

static std::string str(std::vector<int> const& v)
{
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}

static Box3f box(float lo, float hi) { return Box3f{Vec3f{{lo, lo, lo}}, Vec3f{{hi, hi, hi}}}; }

static Mat44f translate(float t)
{
  Mat44f m = Mat44f::identity();
  m[3][0] = m[3][1] = m[3][2] = t;
  return m;
}

static bool same_box(Box3f const& a, Box3f const& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// ---------------------------------------------------------------------------
// 1. hand-traced scenario
// ---------------------------------------------------------------------------
static void scenario(Manager& m, bool refresh, const char* who)
{
  const Mat44f I = Mat44f::identity();
  std::vector<Mat44f> mats(4, I);
  auto g = m.groupCreate();
  m.groupSetMatrices(g, mats.data(), mats.size(), sizeof(Mat44f));
  std::vector<ObjectSharedPtr> o;
  const float lo[4] = {-0.5f, 5.f, -0.2f, 7.f}, hi[4] = {0.5f, 6.f, 0.2f, 8.f};
  for (int i = 0; i < 3; ++i) {
    o.push_back(m.objectCreate(std::make_shared<Id>(i)));
    m.objectSetTransformIndex(o[i], 0);
    m.objectSetBoundingBox(o[i], box(lo[i], hi[i]));
    m.groupAddObject(g, o[i]);
  }
  auto r = m.groupCreateResult(g);
  // objects start visible (ResultBitSet.cpp:74-78): the first cull reports the invisible B
  EXPECT(m.resultObjectIsVisible(r, o[1]), "%s: visible before the first cull", who);
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == std::vector<int>{1}, "%s: first cull changed %s", who,
         str(ids(m, m.resultGetChanged(r))).c_str());
  EXPECT(m.resultObjectIsVisible(r, o[0]) && !m.resultObjectIsVisible(r, o[1]) && m.resultObjectIsVisible(r, o[2]),
         "%s: visibility after the first cull", who);
  m.cull(g, r, I);
  EXPECT(m.resultGetChanged(r).empty(), "%s: nothing changed", who);
  // adding an object already in a group throws (GroupBitSet.cpp:87-90)
  bool threw = false;
  try {
    m.groupAddObject(g, o[0]);
  } catch (std::runtime_error const&) {
    threw = true;
  }
  EXPECT(threw, "%s: double add throws", who);
  // remove A: C moves to index 0 and keeps its bit (ResultBitSet.cpp:110-128)
  m.groupRemoveObject(g, o[0]);
  EXPECT(m.groupGetCount(g) == 2 && id_of(m, m.groupGetObject(g, 0)) == 2, "%s: swap-remove", who);
  EXPECT(m.resultObjectIsVisible(r, o[2]), "%s: C keeps its visible bit at index 0", who);
  threw = false;
  try {
    m.groupRemoveObject(g, o[0]);
  } catch (std::runtime_error const&) {
    threw = true;
  }
  EXPECT(threw, "%s: removing a non-member throws", who);
  m.cull(g, r, I);
  EXPECT(m.resultGetChanged(r).empty(), "%s: after remove nothing changed (%s)", who,
         str(ids(m, m.resultGetChanged(r))).c_str());
  // add D (invisible): the result grows with D visible, the cull flips it
  o.push_back(m.objectCreate(std::make_shared<Id>(3)));
  m.objectSetTransformIndex(o[3], 0);
  m.objectSetBoundingBox(o[3], box(lo[3], hi[3]));
  m.groupAddObject(g, o[3]);
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == std::vector<int>{3}, "%s: D reported (%s)", who,
         str(ids(m, m.resultGetChanged(r))).c_str());
  const Box3f bb0 = m.getBoundingBox(g);
  // move B into view through its box: the CPU cache ignores it until refreshed
  m.objectSetBoundingBox(o[1], box(-0.9f, 0.9f));
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == (refresh ? std::vector<int>{1} : std::vector<int>{}),
         "%s: box change seen=%d (%s)", who, refresh, str(ids(m, m.resultGetChanged(r))).c_str());
  const Box3f bb1 = m.getBoundingBox(g);
  EXPECT(same_box(bb0, bb1) == !refresh, "%s: bounding box cached=%d", who, !refresh);
  // a matrix change notification refreshes the whole cache
  m.groupMatrixChanged(g, 0);
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == (refresh ? std::vector<int>{} : std::vector<int>{1}),
         "%s: after groupMatrixChanged (%s)", who, str(ids(m, m.resultGetChanged(r))).c_str());
  // move D into view through its matrix (index 1, changed in place and notified)
  mats[1] = translate(-7.5f);
  m.groupMatrixChanged(g, 1);
  m.objectSetTransformIndex(o[3], 1);
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == std::vector<int>{3}, "%s: D moved into view (%s)", who,
         str(ids(m, m.resultGetChanged(r))).c_str());
  const Box3f bb2 = m.getBoundingBox(g);
  const float ext = 0.9f - (-0.9f);       // Boxnt::getSize in FP32
  const float up = -0.9f + ext;           // lower corner + extent, as the corner walk adds it
  EXPECT(bb2.lower[0] == -0.9f && bb2.upper[0] == up, "%s: bounding box (%g %g)", who, bb2.lower[0], bb2.upper[0]);
  // a new matrix array re-reads everything
  std::vector<Mat44f> mats2(4, translate(100.f));
  m.groupSetMatrices(g, mats2.data(), mats2.size(), sizeof(Mat44f));
  m.cull(g, r, I);
  EXPECT(ids(m, m.resultGetChanged(r)) == (std::vector<int>{2, 1, 3}), "%s: everything moved away (%s)", who,
         str(ids(m, m.resultGetChanged(r))).c_str());
}

// ---------------------------------------------------------------------------
// 2. random traces
// ---------------------------------------------------------------------------
struct World {
  Manager* m;
  std::vector<GroupSharedPtr> groups;
  std::vector<ResultSharedPtr> results;  // two per group
  std::vector<ObjectSharedPtr> objects;
};

static void random_trace(bool refresh, unsigned seed, int steps, bool via_adapter = false)
{
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  // the HIP manager directly, or behind the reference-interface adapter
  std::unique_ptr<Manager> hipm(via_adapter ? static_cast<Manager*>(new AdapterManager(hip::Options{refresh}))
                                            : hip::Manager::create(hip::Options{refresh}));
  oracle::CpuManager cpum(refresh);
  World w[2] = {{hipm.get(), {}, {}, {}}, {&cpum, {}, {}, {}}};
  const int G = 3, M = 64;
  std::vector<std::vector<Mat44f>> mats(G);
  auto rand_mat = [&]() {
    Mat44f mm = Mat44f::identity();
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) mm[r][c] = (r == c ? 1.f : 0.f) + 0.4f * u(rng);
    for (int c = 0; c < 3; ++c) mm[3][c] = 30.f * u(rng);
    return mm;
  };
  auto rand_box = [&]() {
    const float c[3] = {40.f * u(rng), 40.f * u(rng), 40.f * u(rng)};
    Box3f b;
    for (int k = 0; k < 3; ++k) {
      const float h = 0.1f + 3.f * (u(rng) + 1.f);
      b.lower[k] = c[k] - h;
      b.upper[k] = c[k] + h;
    }
    return b;
  };
  for (int g = 0; g < G; ++g) {
    mats[g].resize(M);
    for (auto& mm : mats[g]) mm = rand_mat();
    for (auto& x : w) {
      x.groups.push_back(x.m->groupCreate());
      x.m->groupSetMatrices(x.groups.back(), mats[g].data(), M, sizeof(Mat44f));
      x.results.push_back(x.m->groupCreateResult(x.groups.back()));
      x.results.push_back(x.m->groupCreateResult(x.groups.back()));
    }
  }
  std::vector<int> member;  // group of each object, -1 = none
  size_t culls = 0, changes = 0;
  for (int step = 0; step < steps; ++step) {
    const int op = rng() % 100;
    if (op < 30 || member.size() < 20) {  // new object into a group
      const int g = rng() % G;
      const Box3f b = rand_box();
      const size_t t = rng() % M;
      for (auto& x : w) {
        auto o = x.m->objectCreate(std::make_shared<Id>(static_cast<int>(member.size())));
        x.m->objectSetBoundingBox(o, b);
        x.m->objectSetTransformIndex(o, t);
        x.m->groupAddObject(x.groups[g], o);
        x.objects.push_back(o);
      }
      member.push_back(g);
    } else if (op < 40) {  // remove
      const size_t k = rng() % member.size();
      if (member[k] >= 0) {
        for (auto& x : w) x.m->groupRemoveObject(x.groups[member[k]], x.objects[k]);
        member[k] = -1;
      }
    } else if (op < 48) {  // re-add a removed object elsewhere
      const size_t k = rng() % member.size();
      if (member[k] < 0) {
        const int g = rng() % G;
        for (auto& x : w) x.m->groupAddObject(x.groups[g], x.objects[k]);
        member[k] = g;
      }
    } else if (op < 56) {  // box change
      const size_t k = rng() % member.size();
      const Box3f b = rand_box();
      for (auto& x : w) x.m->objectSetBoundingBox(x.objects[k], b);
    } else if (op < 60) {  // transform index change
      const size_t k = rng() % member.size();
      const size_t t = rng() % M;
      for (auto& x : w) x.m->objectSetTransformIndex(x.objects[k], t);
    } else if (op < 70) {  // in-place matrix change, notified (intended) or sometimes silent (CPU cache)
      const int g = rng() % G;
      const size_t i = rng() % M;
      mats[g][i] = rand_mat();
      const bool notify = refresh || (rng() % 2);
      if (notify)
        for (auto& x : w) x.m->groupMatrixChanged(x.groups[g], i);
    } else if (op < 72) {  // a new matrix array
      const int g = rng() % G;
      std::vector<Mat44f> fresh(M);
      for (auto& mm : fresh) mm = rand_mat();
      mats[g].swap(fresh);  // the old array stays alive in `fresh` until the end of this step
      for (auto& x : w) x.m->groupSetMatrices(x.groups[g], mats[g].data(), M, sizeof(Mat44f));
      // both managers re-read the new array before `fresh` (the old one) is freed
      for (auto& x : w) x.m->cull(x.groups[g], x.results[2 * g], Mat44f::identity());
    } else if (op < 94) {  // cull one of the results with a random camera
      const int g = rng() % G;
      const int r = 2 * g + (rng() % 2);
      Mat44f vp = Mat44f::identity();
      const float s = 0.02f + 0.05f * (u(rng) + 1.f);
      vp[0][0] = s;
      vp[1][1] = s;
      vp[2][2] = s;
      vp[3][0] = 0.5f * u(rng);
      vp[3][1] = 0.5f * u(rng);
      vp[3][2] = 0.5f * u(rng);
      if (rng() % 3 == 0) {  // a perspective-like w row
        vp[2][3] = 0.01f * u(rng);
        vp[3][3] = 1.f + 0.5f * u(rng);
      }
      std::vector<int> ch[2];
      for (int k = 0; k < 2; ++k) {
        w[k].m->cull(w[k].groups[g], w[k].results[r], vp);
        ch[k] = ids(*w[k].m, w[k].m->resultGetChanged(w[k].results[r]));
      }
      ++culls;
      changes += ch[1].size();
      EXPECT(ch[0] == ch[1], "seed %u step %d: changed lists differ: hip %s oracle %s", seed, step, str(ch[0]).c_str(),
             str(ch[1]).c_str());
      for (size_t k = 0; k < member.size(); ++k) {
        if (member[k] != g) continue;
        EXPECT(w[0].m->resultObjectIsVisible(w[0].results[r], w[0].objects[k]) ==
                   w[1].m->resultObjectIsVisible(w[1].results[r], w[1].objects[k]),
               "seed %u step %d: isVisible of object %zu", seed, step, k);
      }
    } else {  // bounding box
      const int g = rng() % G;
      const Box3f a = w[0].m->getBoundingBox(w[0].groups[g]);
      const Box3f b = w[1].m->getBoundingBox(w[1].groups[g]);
      EXPECT(same_box(a, b), "seed %u step %d: bounding boxes differ (%g..%g vs %g..%g)", seed, step, a.lower[0],
             a.upper[0], b.lower[0], b.upper[0]);
    }
  }
  std::printf("  random trace seed %u (%s%s): %zu culls, %zu changed objects, %zu objects\n", seed,
              refresh ? "refresh" : "cpu-cache", via_adapter ? ", through the adapter" : "", culls, changes,
              member.size());
}

// ---------------------------------------------------------------------------
// 3. cullBatch == sequence of cull()
// ---------------------------------------------------------------------------
static void batch_equals_sequence()
{
  std::mt19937 rng(77);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  std::unique_ptr<Manager> a(hip::Manager::create()), b(hip::Manager::create());
  const int G = 4, M = 32;
  std::vector<std::vector<Mat44f>> mats(G, std::vector<Mat44f>(M, Mat44f::identity()));
  for (auto& v : mats)
    for (auto& mm : v)
      for (int c = 0; c < 3; ++c) mm[3][c] = 20.f * u(rng);
  std::vector<GroupSharedPtr> ga, gb;
  std::vector<ResultSharedPtr> ra, rb;
  for (int g = 0; g < G; ++g) {
    ga.push_back(a->groupCreate());
    gb.push_back(b->groupCreate());
    a->groupSetMatrices(ga[g], mats[g].data(), M, 64);
    b->groupSetMatrices(gb[g], mats[g].data(), M, 64);
    const int n = 100 + 700 * g;
    for (int i = 0; i < n; ++i) {
      Box3f bx = box(0, 0);
      for (int k = 0; k < 3; ++k) {
        bx.lower[k] = 30.f * u(rng);
        bx.upper[k] = bx.lower[k] + 1.f;
      }
      const size_t t = rng() % M;
      auto oa = a->objectCreate(std::make_shared<Id>(i));
      auto ob = b->objectCreate(std::make_shared<Id>(i));
      a->objectSetBoundingBox(oa, bx);
      b->objectSetBoundingBox(ob, bx);
      a->objectSetTransformIndex(oa, t);
      b->objectSetTransformIndex(ob, t);
      a->groupAddObject(ga[g], oa);
      b->groupAddObject(gb[g], ob);
    }
    for (int k = 0; k < 3; ++k) {
      ra.push_back(a->groupCreateResult(ga[g]));
      rb.push_back(b->groupCreateResult(gb[g]));
    }
  }
  for (int frame = 0; frame < 3; ++frame) {
    std::vector<hip::CullRequest> req;
    std::vector<int> which;
    for (int k = 0; k < 10; ++k) {
      const int r = rng() % (3 * G);
      Mat44f vp = Mat44f::identity();
      const float s = 0.03f + 0.03f * (u(rng) + 1.f);
      vp[0][0] = vp[1][1] = vp[2][2] = s;
      vp[3][0] = 0.3f * u(rng);
      req.push_back(hip::CullRequest{gb[r / 3], rb[r], vp});
      which.push_back(r);
    }
    hip::cullBatch(b.get(), req);
    for (size_t k = 0; k < req.size(); ++k) {
      a->cull(ga[which[k] / 3], ra[which[k]], req[k].viewProjection);
      // the batch's result for a request is overwritten by later requests on the same
      // result, so compare final visibility below and changed lists for the last use
      bool last = true;
      for (size_t j = k + 1; j < req.size(); ++j) last = last && which[j] != which[k];
      if (last)
        EXPECT(ids(*a, a->resultGetChanged(ra[which[k]])) == ids(*b, b->resultGetChanged(rb[which[k]])),
               "frame %d request %zu: batch changed list differs", frame, k);
    }
    for (int r = 0; r < 3 * G; ++r)
      for (size_t i = 0; i < a->groupGetCount(ga[r / 3]); ++i)
        EXPECT(a->resultObjectIsVisible(ra[r], a->groupGetObject(ga[r / 3], i)) ==
                   b->resultObjectIsVisible(rb[r], b->groupGetObject(gb[r / 3], i)),
               "frame %d result %d object %zu", frame, r, i);
  }
  std::printf("  cullBatch == cull sequence: checked\n");
}

// ---------------------------------------------------------------------------
// 4. group bound to a GPU tree
// ---------------------------------------------------------------------------
static void tree_bound_group()
{
  std::mt19937 rng(5);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  transform::Tree tree;
  std::vector<transform::Index> nodes{0};
  for (int i = 0; i < 5000; ++i) {
    Mat44f m = Mat44f::identity();
    for (int c = 0; c < 3; ++c) m[3][c] = 4.f * u(rng);
    m[0][1] = 0.2f * u(rng);
    nodes.push_back(tree.addTransform(nodes[rng() % nodes.size()], m));
  }
  tree.compute();
  std::unique_ptr<Manager> hipm(hip::Manager::create());
  oracle::CpuManager cpum(true);
  auto gh = hipm->groupCreate();
  auto gc = cpum.groupCreate();
  hip::groupSetTreeMatrices(hipm.get(), gh, tree);
  cpum.groupSetMatrices(gc, tree.getWorldMatrices(), tree.getTransformCount(), sizeof(Mat44f));
  for (int i = 0; i < 20000; ++i) {
    Box3f bx;
    for (int k = 0; k < 3; ++k) {
      bx.lower[k] = 0.5f * u(rng);
      bx.upper[k] = bx.lower[k] + 0.3f;
    }
    const size_t t = nodes[1 + rng() % (nodes.size() - 1)];
    auto oh = hipm->objectCreate(std::make_shared<Id>(i));
    auto oc = cpum.objectCreate(std::make_shared<Id>(i));
    hipm->objectSetBoundingBox(oh, bx);
    cpum.objectSetBoundingBox(oc, bx);
    hipm->objectSetTransformIndex(oh, t);
    cpum.objectSetTransformIndex(oc, t);
    hipm->groupAddObject(gh, oh);
    cpum.groupAddObject(gc, oc);
  }
  auto rh = hipm->groupCreateResult(gh);
  auto rc = cpum.groupCreateResult(gc);
  for (int frame = 0; frame < 3; ++frame) {
    if (frame) {
      for (int k = 0; k < 300; ++k) {
        Mat44f m = Mat44f::identity();
        for (int c = 0; c < 3; ++c) m[3][c] = 4.f * u(rng);
        tree.updateLocalMatrix(nodes[1 + rng() % (nodes.size() - 1)], m);
      }
      tree.compute();
      // the CPU side reads the host copy of the world matrices; everything may have moved
      cpum.groupSetMatrices(gc, tree.getWorldMatrices(), tree.getTransformCount(), sizeof(Mat44f));
      for (size_t i = 0; i < tree.getTransformCount(); ++i) cpum.groupMatrixChanged(gc, i);
    }
    Mat44f vp = Mat44f::identity();
    vp[0][0] = vp[1][1] = vp[2][2] = 0.05f;
    hipm->cull(gh, rh, vp);
    cpum.cull(gc, rc, vp);
    const auto a = ids(*hipm, hipm->resultGetChanged(rh));
    const auto b = ids(cpum, cpum.resultGetChanged(rc));
    EXPECT(a == b, "tree frame %d: changed lists differ (%zu vs %zu)", frame, a.size(), b.size());
    const Box3f ba = hipm->getBoundingBox(gh), bb = cpum.getBoundingBox(gc);
    EXPECT(same_box(ba, bb), "tree frame %d: bounding boxes differ", frame);
    std::printf("  tree-bound group frame %d: %zu changed\n", frame, a.size());
  }
}

int main()
{
  {
    oracle::CpuManager cpu_cache(false), cpu_refresh(true);
    scenario(cpu_cache, false, "oracle/cpu-cache");
    scenario(cpu_refresh, true, "oracle/refresh");
  }
  {
    std::unique_ptr<Manager> a(hip::Manager::create(hip::Options{false}));
    std::unique_ptr<Manager> b(hip::Manager::create(hip::Options{true}));
    scenario(*a, false, "hip/cpu-cache");
    scenario(*b, true, "hip/refresh");
  }
  {
    // the same hand-traced expectations through the reference-interface adapter
    AdapterManager a(hip::Options{false}), b(hip::Options{true});
    scenario(a, false, "adapter/cpu-cache");
    scenario(b, true, "adapter/refresh");
  }
  for (unsigned seed : {1u, 2u, 3u}) {
    random_trace(false, seed, 3000);
    random_trace(true, seed + 100, 3000);
  }
  random_trace(false, 7, 3000, true);
  random_trace(true, 107, 3000, true);
  batch_equals_sequence();
  tree_bound_group();
  std::printf("manager parity: %d failures\n", failures);
  return failures ? 1 : 0;
}

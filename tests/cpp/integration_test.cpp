// integration_test.cpp -- the xbar side of the integration, on the GPU.
//
// XbarCulling below makes exactly the calls xbar::culling::CullingImpl makes
// (dp/sg/xbar/culling/src/CullingImpl.cpp):
//   construction   groupCreate, then addObject for every drawable (:52-101)
//   addObject      objectCreate(Payload(objectTreeIndex)), groupAddObject,
//                  objectSetTransformIndex(node.m_transform), objectSetBoundingBox (:158-178)
//   ADDED/REMOVED/CHANGED events   addObject / groupRemoveObject / objectSetBoundingBox (:180-205)
//   cull           groupSetMatrices(tree.getWorldMatrices(), tree.getTransformCount(), 64),
//                  cull, then every changed object mapped to its ObjectTreeIndex through
//                  objectGetUserData (:126-144)
//   getBoundingBox groupSetMatrices as above, then getBoundingBox (:146-151)
// plus the TransformObserver the reference declares but never attaches
// (CullingImpl.h:99-111, forwarding body :216-223), here
// dpcull::adapter::TransformForwarder attached to the transform tree.
//
// The manager behind it is the reference-interface adapter over the HIP
// backend (integration/dpcull_adapter.hpp) on one side and the CPU manager
// restatement (oracle/cpu_manager.hpp) on the other, both fed by one GPU
// transform tree (dpcull::transform::Tree) whose local matrices move every
// frame.  Checked per frame and per camera: the changed ObjectTreeIndex lists
// are equal, every object's visibility is equal, the bounding boxes are equal
// bit for bit; the forwarded results equal a full cull of the current world
// matrices by the oracle; without the forwarder the HIP results go stale
// (the observer matters); a group bound to the tree's device matrices gives
// the same lists.
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "cpu_manager.hpp"
#include "dpcull/Manager.hpp"
#include "dpcull/Tree.hpp"
#include "oracle.h"
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

typedef uint32_t ObjectTreeIndex;

// CullingImpl::Payload (CullingImpl.h:73-97): the ObjectTreeIndex of a culling object.
struct TreeIndexPayload : Payload {
  explicit TreeIndexPayload(ObjectTreeIndex i) : index(i) {}
  ObjectTreeIndex index;
};

// The scene side xbar reads: per ObjectTreeIndex, a drawable's transform index
// (ObjectTreeNode::m_transform, ObjectTree.h:82) and its bounding box.
struct SceneObjects {
  std::vector<transform::Index> transform;
  std::vector<Box3f> box;
  std::vector<bool> present;
};

class XbarCulling {
 public:
  enum class Matrices { HostArray, DeviceTree };

  XbarCulling(Manager* manager, transform::Tree& tree, SceneObjects const& scene, Matrices mode, bool forward)
      : m_culling(manager), m_tree(tree), m_scene(scene), m_mode(mode)
  {
    m_group = m_culling->groupCreate();
    if (mode == Matrices::DeviceTree) {
      if (auto* a = dynamic_cast<AdapterManager*>(m_culling.get())) a->groupSetTreeMatrices(m_group, m_tree);
    }
    if (forward) {
      m_forwarder.reset(new adapter::TransformForwarder<Manager, GroupSharedPtr>(*m_culling, m_group));
      m_tree.attach(m_forwarder.get());
    }
    for (ObjectTreeIndex i = 0; i < scene.present.size(); ++i)
      if (scene.present[i]) addObject(i);
  }
  ~XbarCulling()
  {
    if (m_forwarder) m_tree.detach(m_forwarder.get());
  }

  ResultSharedPtr resultCreate() { return m_culling->groupCreateResult(m_group); }

  bool resultIsVisible(ResultSharedPtr const& r, ObjectTreeIndex i) const
  {
    return m_culling->resultObjectIsVisible(r, m_objects[i]);
  }

  std::vector<ObjectTreeIndex> cull(ResultSharedPtr const& result, Mat44f const& vp)
  {
    setMatrices();
    m_culling->cull(m_group, result, vp);
    std::vector<ObjectTreeIndex> changed;
    for (auto const& o : m_culling->resultGetChanged(result))
      changed.push_back(std::static_pointer_cast<TreeIndexPayload>(m_culling->objectGetUserData(o))->index);
    return changed;
  }

  Box3f getBoundingBox()
  {
    setMatrices();
    return m_culling->getBoundingBox(m_group);
  }

  // SceneTree events
  void added(ObjectTreeIndex i) { addObject(i); }
  void removed(ObjectTreeIndex i)
  {
    m_culling->groupRemoveObject(m_group, m_objects[i]);
    m_objects[i].reset();
  }
  void changed(ObjectTreeIndex i) { m_culling->objectSetBoundingBox(m_objects[i], m_scene.box[i]); }

 private:
  void setMatrices()
  {
    if (m_mode == Matrices::HostArray)
      m_culling->groupSetMatrices(m_group, m_tree.getWorldMatrices(), m_tree.getTransformCount(), sizeof(Mat44f));
  }
  void addObject(ObjectTreeIndex i)
  {
    if (m_objects.size() != m_scene.present.size()) m_objects.resize(m_scene.present.size());
    m_objects[i] = m_culling->objectCreate(std::make_shared<TreeIndexPayload>(i));
    m_culling->groupAddObject(m_group, m_objects[i]);
    m_culling->objectSetTransformIndex(m_objects[i], m_scene.transform[i]);
    m_culling->objectSetBoundingBox(m_objects[i], m_scene.box[i]);
  }

  std::unique_ptr<Manager> m_culling;
  transform::Tree& m_tree;
  SceneObjects const& m_scene;
  Matrices m_mode;
  GroupSharedPtr m_group;
  std::vector<ObjectSharedPtr> m_objects;
  std::unique_ptr<transform::TreeObserver> m_forwarder;
};

static std::string head(std::vector<ObjectTreeIndex> const& v)
{
  std::string s = "[";
  for (size_t i = 0; i < v.size() && i < 8; ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + (v.size() > 8 ? ",...]" : "]");
}

static Mat44f perspective_like(std::mt19937& rng)
{
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  Mat44f vp = Mat44f::identity();
  const float s = 0.03f + 0.02f * (u(rng) + 1.f);
  vp[0][0] = s;
  vp[1][1] = s * 1.5f;
  vp[2][2] = s;
  vp[2][3] = 0.004f * u(rng);
  vp[3][0] = 0.3f * u(rng);
  vp[3][1] = 0.3f * u(rng);
  vp[3][3] = 1.f + 0.2f * u(rng);
  return vp;
}

static void xbar_frames(unsigned seed)
{
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  auto rand_local = [&](float t) {
    Mat44f m = Mat44f::identity();
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r][c] = (r == c ? 1.f : 0.f) + 0.25f * u(rng);
    for (int c = 0; c < 3; ++c) m[3][c] = t * u(rng);
    return m;
  };
  transform::Tree tree;
  std::vector<transform::Index> nodes{0};
  for (int i = 0; i < 3000; ++i) nodes.push_back(tree.addTransform(nodes[rng() % nodes.size()], rand_local(6.f)));
  tree.compute();

  SceneObjects scene;
  const size_t n_objects = 6000;
  auto rand_box = [&]() {
    Box3f b;
    for (int k = 0; k < 3; ++k) {
      b.lower[k] = 2.f * u(rng);
      b.upper[k] = b.lower[k] + 0.1f + 0.8f * (u(rng) + 1.f);
    }
    return b;
  };
  for (size_t i = 0; i < n_objects; ++i) {
    scene.transform.push_back(nodes[1 + rng() % (nodes.size() - 1)]);
    scene.box.push_back(rand_box());
    scene.present.push_back(rng() % 10 != 0);  // some ObjectTree indices are not drawables yet
  }

  // the reference's caching (Options{false} / CpuManager(false)), host matrices, forwarded tree events
  XbarCulling hip_fwd(new AdapterManager(hip::Options{false}), tree, scene, XbarCulling::Matrices::HostArray, true);
  XbarCulling cpu_fwd(new oracle::CpuManager(false), tree, scene, XbarCulling::Matrices::HostArray, true);
  // no forwarder: groupSetMatrices with the same pointer changes nothing, matrices go stale
  XbarCulling hip_stale(new AdapterManager(hip::Options{false}), tree, scene, XbarCulling::Matrices::HostArray, false);
  // the tree's device matrices (no host copy); box changes apply at once (Options{true})
  XbarCulling hip_tree(new AdapterManager(hip::Options{true}), tree, scene, XbarCulling::Matrices::DeviceTree, false);
  XbarCulling cpu_refresh(new oracle::CpuManager(true), tree, scene, XbarCulling::Matrices::HostArray, true);

  const int cameras = 2;
  std::vector<ResultSharedPtr> r_hip, r_cpu, r_stale, r_tree, r_ref;
  for (int c = 0; c < cameras; ++c) {
    r_hip.push_back(hip_fwd.resultCreate());
    r_cpu.push_back(cpu_fwd.resultCreate());
    r_stale.push_back(hip_stale.resultCreate());
    r_tree.push_back(hip_tree.resultCreate());
    r_ref.push_back(cpu_refresh.resultCreate());
  }
  size_t stale_frames = 0, total_changed = 0;
  for (int frame = 0; frame < 8; ++frame) {
    if (frame) {
      // transforms move (dirty local matrices), then TransformTree::compute
      for (int k = 0; k < 150; ++k) tree.updateLocalMatrix(nodes[1 + rng() % (nodes.size() - 1)], rand_local(6.f));
      // scene events on even frames: drawables appear and disappear, boxes
      // change (an add or remove refreshes every group's cache, so odd frames,
      // with moved transforms only, show what the forwarder adds)
      for (int k = 0; k < (frame % 2 ? 0 : 40); ++k) {
        const ObjectTreeIndex i = static_cast<ObjectTreeIndex>(rng() % n_objects);
        if (!scene.present[i]) {
          scene.present[i] = true;
          for (XbarCulling* x : {&hip_fwd, &cpu_fwd, &hip_stale, &hip_tree, &cpu_refresh}) x->added(i);
        } else if (rng() % 2) {
          scene.present[i] = false;
          for (XbarCulling* x : {&hip_fwd, &cpu_fwd, &hip_stale, &hip_tree, &cpu_refresh}) x->removed(i);
        } else {
          scene.box[i] = rand_box();
          for (XbarCulling* x : {&hip_fwd, &cpu_fwd, &hip_stale, &hip_tree, &cpu_refresh}) x->changed(i);
        }
      }
      tree.compute();
    }
    for (int c = 0; c < cameras; ++c) {
      const Mat44f vp = perspective_like(rng);
      const auto a = hip_fwd.cull(r_hip[c], vp);
      const auto b = cpu_fwd.cull(r_cpu[c], vp);
      const auto s = hip_stale.cull(r_stale[c], vp);
      const auto t = hip_tree.cull(r_tree[c], vp);
      const auto d = cpu_refresh.cull(r_ref[c], vp);
      total_changed += a.size();
      EXPECT(a == b, "seed %u frame %d camera %d: changed indices differ: hip %s oracle %s", seed, frame, c,
             head(a).c_str(), head(b).c_str());
      EXPECT(t == d, "seed %u frame %d camera %d: tree-bound changed indices differ: hip %s oracle %s", seed, frame, c,
             head(t).c_str(), head(d).c_str());
      // visibility of every drawable, and against a full cull of the current world matrices
      const Mat44f* w = tree.getWorldMatrices();
      size_t bad = 0, bad_fresh = 0, stale = 0, bad_tree = 0;
      for (ObjectTreeIndex i = 0; i < n_objects; ++i) {
        if (!scene.present[i]) continue;
        const bool va = hip_fwd.resultIsVisible(r_hip[c], i);
        bad += va != cpu_fwd.resultIsVisible(r_cpu[c], i);
        bad_tree += hip_tree.resultIsVisible(r_tree[c], i) != cpu_refresh.resultIsVisible(r_ref[c], i);
        stale += va != hip_stale.resultIsVisible(r_stale[c], i);
        // the oracle's cull of this object now: box, world matrix of its transform, the camera
        const Box3f& bx = scene.box[i];
        const float l4[4] = {bx.lower[0], bx.lower[1], bx.lower[2], 1.f};
        const float e4[4] = {bx.upper[0] - bx.lower[0], bx.upper[1] - bx.lower[1], bx.upper[2] - bx.lower[2], 0.f};
        const uint64_t xf = scene.transform[i];
        uint32_t bit[2] = {0, 0};
        orc_cull_group(1, l4, e4, &xf, w, sizeof(Mat44f), vp.m, bit);
        bad_fresh += va != ((bit[0] & 1u) != 0);
      }
      EXPECT(bad == 0, "seed %u frame %d camera %d: %zu visibilities differ from the oracle manager", seed, frame, c, bad);
      EXPECT(bad_tree == 0, "seed %u frame %d camera %d: %zu tree-bound visibilities differ", seed, frame, c, bad_tree);
      // every frame moves transforms, so the forwarded group re-reads the moved
      // matrices and with them refreshes its cache: the result is current
      EXPECT(bad_fresh == 0, "seed %u frame %d camera %d: %zu visibilities differ from a cull of the current scene",
             seed, frame, c, bad_fresh);
      if (frame % 2 && stale) ++stale_frames;
    }
    const Box3f ba = hip_fwd.getBoundingBox(), bb = cpu_fwd.getBoundingBox();
    EXPECT(std::memcmp(&ba, &bb, sizeof ba) == 0, "seed %u frame %d: bounding boxes differ", seed, frame);
    const Box3f ta = hip_tree.getBoundingBox(), tb = cpu_refresh.getBoundingBox();
    EXPECT(std::memcmp(&ta, &tb, sizeof ta) == 0, "seed %u frame %d: tree-bound bounding boxes differ", seed, frame);
  }
  // without the forwarder the results lag behind the moving transforms
  EXPECT(stale_frames > 0, "seed %u: a group without the tree forwarder never went stale", seed);
  std::printf("  xbar frames seed %u: 8 frames x %d cameras, %zu changed indices, stale without forwarding in %zu\n",
              seed, cameras, total_changed, stale_frames);
}

// The forwarder reports exactly the tree's changed-matrix bits, each once, in
// ascending order, as CullingImpl::TransformObserver's traverseBits loop.
struct CountingManager : oracle::CpuManager {
  std::vector<size_t> calls;
  void groupMatrixChanged(GroupSharedPtr const& g, size_t i) override
  {
    calls.push_back(i);
    oracle::CpuManager::groupMatrixChanged(g, i);
  }
};

static void forwarder_reports_event_bits()
{
  transform::Tree tree;
  std::vector<transform::Index> nodes{0};
  std::mt19937 rng(9);
  for (int i = 0; i < 500; ++i) nodes.push_back(tree.addTransform(nodes[rng() % nodes.size()], Mat44f::identity()));
  CountingManager m;
  auto g = m.groupCreate();
  m.groupSetMatrices(g, tree.getWorldMatrices(), tree.getTransformCount(), sizeof(Mat44f));
  adapter::TransformForwarder<CountingManager, GroupSharedPtr> fwd(m, g);
  tree.attach(&fwd);
  for (int frame = 0; frame < 3; ++frame) {
    if (frame)
      for (int k = 0; k < 20; ++k) {
        Mat44f mm = Mat44f::identity();
        mm[3][0] = static_cast<float>(k + frame);
        tree.updateLocalMatrix(nodes[1 + rng() % (nodes.size() - 1)], mm);
      }
    m.calls.clear();
    tree.compute();
    std::vector<size_t> want;
    const auto& bits = tree.getLastChangedWorldMatrices();
    for (size_t i = 0; i < tree.getTransformCount(); ++i)
      if ((bits[i >> 6] >> (i & 63)) & 1) want.push_back(i);
    EXPECT(m.calls == want, "frame %d: forwarded %zu indices, the event holds %zu", frame, m.calls.size(), want.size());
    EXPECT(!want.empty(), "frame %d: the event is empty", frame);
  }
  tree.detach(&fwd);
  std::printf("  forwarder: groupMatrixChanged once per event bit, ascending\n");
}

int main()
{
  forwarder_reports_event_bits();
  xbar_frames(1);
  xbar_frames(2);
  std::printf("integration: %d failures\n", failures);
  return failures ? 1 : 0;
}

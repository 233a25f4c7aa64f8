// Manager.cpp -- the HIP culling manager behind include/dpcull/Manager.hpp.
//
// Host bookkeeping follows dp/culling/src/{ManagerBitSet,GroupBitSet,ObjectBitSet,
// ResultBitSet}.cpp; every per-object computation (OBB, 8-corner test, result
// diff, bounding box) runs on the device through the C ABI.
#include "dpcull/Manager.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "dpcull/Tree.hpp"
#include "dpcull_hip.h"
#include "trace.hpp"
#include "unwind.hpp"

namespace dpcull {
namespace culling {
namespace hip {

namespace {

void check(int rc, const char* what)
{
  if (rc != DPC_OK) throw std::runtime_error(std::string(what) + ": " + dpc_last_error());
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void reserve(size_t count)
  {
    if (count <= n) return;
    release();
    void* q = nullptr;
    check(dpc_malloc(&q, count * sizeof(T)), "dpc_malloc");
    p = static_cast<T*>(q);
    n = count;
  }
  void release()
  {
    if (p) dpc_free(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

const size_t kNone = ~size_t(0);

class GroupImpl;

// ObjectBitSet (dp/culling/ObjectBitSet.h:42-76): lowerLeft = (lower, 1),
// extent = (upper - lower, 0), transform index, payload, position in its group.
class ObjectImpl : public Object {
 public:
  explicit ObjectImpl(PayloadSharedPtr const& userData) : m_userData(userData) {}
  Vec4f m_lowerLeft{{0, 0, 0, 1}};
  Vec4f m_extent{{0, 0, 0, 0}};
  size_t m_transformIndex = kNone;
  PayloadSharedPtr m_userData;
  size_t m_groupIndex = kNone;
  std::weak_ptr<GroupImpl> m_group;
};

class ResultImpl;

// GroupBitSet (dp/culling/GroupBitSet.h:40-104) plus the device arrays of the group.
class GroupImpl : public Group, public std::enable_shared_from_this<GroupImpl> {
 public:
  ~GroupImpl() override
  {
    for (auto& o : m_objects) o->m_groupIndex = kNone;  // GroupBitSet.cpp:67-74
  }

  std::vector<std::shared_ptr<ObjectImpl>> m_objects;
  size_t m_incarnation = 0;
  // matrices borrowed from the caller (GroupBitSet.cpp:121-135)
  const void* m_matrices = nullptr;
  size_t m_matricesCount = 0;
  size_t m_matricesStride = 0;
  std::vector<uint8_t> m_dirtyMatrices;
  bool m_matricesChanged = true;
  bool m_obbDirty = true;           // the CPU backend's OBB-cache flag
  size_t m_incarnationOBB = kNone;  // GroupCPU::m_objectIncarnationOBB
  bool m_objectsDirty = true;       // device object arrays out of date
  bool m_boundingBoxDirty = true;
  uint64_t m_boundingBoxTreeCompute = 0;
  Box3f m_boundingBox;
  std::vector<ResultImpl*> m_results;
  // matrices from a GPU tree instead of the host pointer
  const transform::Tree* m_tree = nullptr;
  // device arrays
  DevBuf<float> box_a, box_b, planes, aos, delta;
  DevBuf<uint32_t> xform, xform_slot, idx;
  DevBuf<uint64_t> mask;
  DevBuf<unsigned char> bbox_ws;
  DevBuf<float> bbox_out;
  size_t deviceObjects = 0;
};

// ResultBitSet (dp/culling/ResultBitSet.h:38-67): last visibility bits (host,
// authoritative for queries and remap events) and their device mirror.
class ResultImpl : public Result {
 public:
  explicit ResultImpl(std::shared_ptr<GroupImpl> const& g) : m_group(g) { g->m_results.push_back(this); }
  ~ResultImpl() override
  {
    auto& r = m_group->m_results;
    r.erase(std::remove(r.begin(), r.end(), this), r.end());
  }

  bool getBit(size_t i) const { return (m_bits[i >> 6] >> (i & 63)) & 1; }
  void setBit(size_t i, bool v)
  {
    if (v)
      m_bits[i >> 6] |= uint64_t(1) << (i & 63);
    else
      m_bits[i >> 6] &= ~(uint64_t(1) << (i & 63));
  }
  // BitArray::resize(n, false): keeps bits < min(old, n), clears the rest.
  void resize(size_t n)
  {
    m_bits.resize((n + 63) / 64, 0);
    if (n % 64 && !m_bits.empty()) m_bits.back() &= (uint64_t(1) << (n % 64)) - 1;
    m_size = n;
  }

  // ResultBitSet::onNotify (ResultBitSet.cpp:110-128): the object at oldIndex moved to newIndex.
  void onRemap(size_t oldIndex, size_t newIndex)
  {
    if (newIndex < m_size) {
      setBit(newIndex, oldIndex < m_size ? getBit(oldIndex) : true);
      m_deviceStale = true;
    }
  }

  std::shared_ptr<GroupImpl> m_group;
  std::vector<ObjectSharedPtr> m_changed;
  size_t m_incarnation = kNone;
  std::vector<uint64_t> m_bits;
  size_t m_size = 0;
  bool m_deviceStale = true;
  DevBuf<uint64_t> d_old;
  DevBuf<uint32_t> d_changed, d_count;
  DevBuf<unsigned char> d_ws;
};

class ManagerImpl : public culling::Manager {
 public:
  explicit ManagerImpl(Options const& o) : m_options(o) { check(dpc_stream_create(&m_stream), "dpc_stream_create"); }
  ~ManagerImpl() override
  {
    if (m_stream) dpc_stream_destroy(m_stream);
  }

  ObjectSharedPtr objectCreate(PayloadSharedPtr const& userData) override
  {
    return std::make_shared<ObjectImpl>(userData);
  }

  void objectSetBoundingBox(ObjectSharedPtr const& object, Box3f const& box) override
  {
    auto o = impl(object);
    // ManagerBitSet.cpp:102-103, FP32 subtraction as Boxnt::getSize (Boxnt.h:239-242).
    o->m_lowerLeft = Vec4f{{box.lower[0], box.lower[1], box.lower[2], 1.0f}};
    o->m_extent = Vec4f{{box.upper[0] - box.lower[0], box.upper[1] - box.lower[1], box.upper[2] - box.lower[2], 0.0f}};
    touched(o);
  }

  void objectSetTransformIndex(ObjectSharedPtr const& object, size_t index) override
  {
    auto o = impl(object);
    o->m_transformIndex = index;
    touched(o);
  }

  void objectSetUserData(ObjectSharedPtr const& object, PayloadSharedPtr const& userData) override
  {
    impl(object)->m_userData = userData;
  }

  PayloadSharedPtr const& objectGetUserData(ObjectSharedPtr const& object) override
  {
    return impl(object)->m_userData;
  }

  GroupSharedPtr groupCreate() override { return std::make_shared<GroupImpl>(); }

  // GroupBitSet::addObject (GroupBitSet.cpp:76-91)
  void groupAddObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) override
  {
    auto g = impl(group);
    auto o = impl(object);
    if (o->m_groupIndex != kNone) throw std::runtime_error("object already belongs to a different group");
    o->m_groupIndex = g->m_objects.size();
    o->m_group = g;
    g->m_objects.push_back(o);
    ++g->m_incarnation;
    g->m_boundingBoxDirty = true;
    g->m_obbDirty = true;
    g->m_objectsDirty = true;
  }

  ObjectSharedPtr groupGetObject(GroupSharedPtr const& group, size_t index) override
  {
    auto g = impl(group);
    if (index >= g->m_objects.size()) throw std::out_of_range("groupGetObject: index beyond the group");
    return g->m_objects[index];
  }

  // GroupBitSet::removeObject (GroupBitSet.cpp:93-119): swap-remove and remap event.
  void groupRemoveObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) override
  {
    auto g = impl(group);
    auto o = impl(object);
    const size_t old = o->m_groupIndex;
    if (!(old < g->m_objects.size() && g->m_objects[old] == o))
      throw std::runtime_error("object does not belong to this group");
    g->m_objects[old] = g->m_objects.back();
    g->m_objects[old]->m_groupIndex = old;
    o->m_groupIndex = kNone;
    o->m_group.reset();
    g->m_objects.pop_back();
    for (ResultImpl* r : g->m_results) r->onRemap(g->m_objects.size(), old);
    g->m_boundingBoxDirty = true;
    g->m_obbDirty = true;
    g->m_objectsDirty = true;
    ++g->m_incarnation;
  }

  size_t groupGetCount(GroupSharedPtr const& group) override { return impl(group)->m_objects.size(); }

  void groupSetMatrices(GroupSharedPtr const& group, void const* matrices, size_t count, size_t stride) override
  {
    auto g = impl(group);
    if (g->m_tree || matrices != g->m_matrices || count != g->m_matricesCount || stride != g->m_matricesStride) {
      g->m_tree = nullptr;
      g->m_matrices = matrices;
      g->m_matricesCount = count;
      g->m_matricesStride = stride;
      g->m_matricesChanged = true;
      g->m_dirtyMatrices.assign(count, 1);
      g->m_boundingBoxDirty = true;
      g->m_obbDirty = true;
      g->m_objectsDirty = true;  // transform indices are re-validated against the new count
    }
  }

  // GroupBitSet::markMatrixDirty (GroupBitSet.h:140-150)
  void groupMatrixChanged(GroupSharedPtr const& group, size_t index) override
  {
    auto g = impl(group);
    if (index < g->m_matricesCount) {
      g->m_dirtyMatrices[index] = 1;
      g->m_boundingBoxDirty = true;
      g->m_obbDirty = true;
    }
  }

  ResultSharedPtr groupCreateResult(GroupSharedPtr const& group) override
  {
    return std::make_shared<ResultImpl>(impl(group));
  }

  std::vector<ObjectSharedPtr> const& resultGetChanged(ResultSharedPtr const& result) override
  {
    return impl(result)->m_changed;
  }

  // ResultBitSet::isVisible (ResultBitSet.h:69-76)
  bool resultObjectIsVisible(ResultSharedPtr const& result, ObjectSharedPtr const& object) override
  {
    auto r = impl(result);
    const size_t i = impl(object)->m_groupIndex;
    return i < r->m_size ? r->getBit(i) : true;
  }

  void cull(GroupSharedPtr const& group, ResultSharedPtr const& result, Mat44f const& viewProjection) override
  {
    TraceRange trace("cull");
    auto g = impl(group);
    auto r = impl(result);
    if (r->m_group != g) throw std::invalid_argument("cull: the result belongs to another group");
    prepare(*g);
    const size_t n = g->m_objects.size();
    const size_t words = (n + 63) / 64;
    g->mask.reserve(std::max<size_t>(words, 1));
    const float* planes = g->m_tree ? g->m_tree->deviceWorldPlanes() : g->planes.p;
    const size_t ld = g->m_tree ? g->m_tree->deviceLd() : g->m_matricesCount;
    const uint32_t* xf = g->m_tree ? g->xform_slot.p : g->xform.p;
    // the cull and the diff of its mask against the result's bits in one call
    // (the cull counts each word's changed bits for the diff)
    updateResult(*g, *r, g->mask.p, n, [&](ResultImpl& res) {
      check(dpc_cull_diff(g->box_a.p, g->box_b.p, xf, n, planes, ld, viewProjection.m, g->mask.p, res.d_old.p,
                          res.d_changed.p, res.d_count.p, 1, res.d_ws.p, m_stream),
            "dpc_cull_diff");
    });
  }

  Box3f getBoundingBox(GroupSharedPtr const& group) const override
  {
    TraceRange trace("cull::getBoundingBox");
    auto g = const_cast<ManagerImpl*>(this)->impl(group);
    // a tree-bound group's world matrices change with every Tree::compute()
    if (g->m_tree && g->m_boundingBoxTreeCompute != g->m_tree->computeCount()) g->m_boundingBoxDirty = true;
    if (g->m_boundingBoxDirty) {
      const_cast<ManagerImpl*>(this)->computeBoundingBox(*g);
      g->m_boundingBoxDirty = false;
      if (g->m_tree) g->m_boundingBoxTreeCompute = g->m_tree->computeCount();
    }
    return g->m_boundingBox;
  }

  // ---- extensions ----------------------------------------------------------
  void bindTree(GroupSharedPtr const& group, transform::Tree const& tree)
  {
    auto g = impl(group);
    g->m_tree = &tree;
    g->m_matrices = nullptr;
    g->m_matricesCount = tree.getTransformCount();
    g->m_matricesStride = sizeof(Mat44f);
    g->m_dirtyMatrices.assign(g->m_matricesCount, 0);
    g->m_matricesChanged = false;
    g->m_objectsDirty = true;  // remap transform indices to tree slots
    g->m_obbDirty = true;
    g->m_boundingBoxDirty = true;
  }

  void cullBatch(std::vector<CullRequest> const& requests);

 private:
  std::shared_ptr<ObjectImpl> impl(ObjectSharedPtr const& o)
  {
    auto p = std::dynamic_pointer_cast<ObjectImpl>(o);
    if (!p) throw std::invalid_argument("not an object of the HIP culling manager");
    return p;
  }
  std::shared_ptr<GroupImpl> impl(GroupSharedPtr const& g)
  {
    auto p = std::dynamic_pointer_cast<GroupImpl>(g);
    if (!p) throw std::invalid_argument("not a group of the HIP culling manager");
    return p;
  }
  std::shared_ptr<ResultImpl> impl(ResultSharedPtr const& r)
  {
    auto p = std::dynamic_pointer_cast<ResultImpl>(r);
    if (!p) throw std::invalid_argument("not a result of the HIP culling manager");
    return p;
  }

  // An object's box or transform index changed.  The reference never links an
  // object to its group, so such changes reach the CPU backend's OBB cache only
  // through a later refresh; with refreshOnObjectChange they are applied now.
  void touched(std::shared_ptr<ObjectImpl> const& o)
  {
    if (!m_options.refreshOnObjectChange) return;
    if (auto g = o->m_group.lock()) {
      g->m_objectsDirty = true;
      g->m_boundingBoxDirty = true;
    }
  }

  void uploadObjects(GroupImpl& g, DevBuf<float>& box_a, DevBuf<float>& box_b, DevBuf<uint32_t>& xform)
  {
    const size_t n = g.m_objects.size();
    std::vector<float> a(4 * n), b(2 * n);
    std::vector<uint32_t> x(n);
    for (size_t i = 0; i < n; ++i) {
      const ObjectImpl& o = *g.m_objects[i];
      a[4 * i + 0] = o.m_lowerLeft[0];
      a[4 * i + 1] = o.m_lowerLeft[1];
      a[4 * i + 2] = o.m_lowerLeft[2];
      a[4 * i + 3] = o.m_extent[0];
      b[2 * i + 0] = o.m_extent[1];
      b[2 * i + 1] = o.m_extent[2];
      // indices outside the matrix array are never dereferenced by the kernels
      x[i] = o.m_transformIndex < g.m_matricesCount ? static_cast<uint32_t>(o.m_transformIndex) : 0xFFFFFFFFu;
    }
    box_a.reserve(std::max<size_t>(4 * n, 4));
    box_b.reserve(std::max<size_t>(2 * n, 2));
    xform.reserve(std::max<size_t>(n, 1));
    SyncOnUnwind sync(m_stream);  // (a, b, x are read by the copies)
    if (n) {
      check(dpc_memcpy_h2d(box_a.p, a.data(), a.size() * 4, m_stream), "upload boxes");
      check(dpc_memcpy_h2d(box_b.p, b.data(), b.size() * 4, m_stream), "upload boxes");
      check(dpc_memcpy_h2d(xform.p, x.data(), x.size() * 4, m_stream), "upload indices");
    }
    check(dpc_stream_sync(m_stream), "upload sync");
  }

  void uploadAllMatrices(GroupImpl& g, DevBuf<float>& planes)
  {
    const size_t count = g.m_matricesCount;
    planes.reserve(std::max<size_t>(16 * count, 16));
    if (!count) return;
    if (!g.m_matrices) throw std::runtime_error("cull: group has no matrices (groupSetMatrices)");
    g.aos.reserve(16 * count);
    // the caller's array may have any stride: pack it, then transpose on the device
    std::vector<float> packed(16 * count);
    const char* base = static_cast<const char*>(g.m_matrices);
    for (size_t k = 0; k < count; ++k) std::memcpy(&packed[16 * k], base + k * g.m_matricesStride, 64);
    SyncOnUnwind sync(m_stream);
    check(dpc_memcpy_h2d(g.aos.p, packed.data(), packed.size() * 4, m_stream), "upload matrices");
    check(dpc_matrices_from_aos(g.aos.p, 64, count, planes.p, count, m_stream), "matrices to planes");
    check(dpc_stream_sync(m_stream), "upload sync");
  }

  void uploadDirtyMatrices(GroupImpl& g)
  {
    std::vector<uint32_t> idx;
    std::vector<float> mats;
    const char* base = static_cast<const char*>(g.m_matrices);
    for (size_t k = 0; k < g.m_matricesCount; ++k) {
      if (!g.m_dirtyMatrices[k]) continue;
      idx.push_back(static_cast<uint32_t>(k));
      mats.resize(mats.size() + 16);
      std::memcpy(&mats[mats.size() - 16], base + k * g.m_matricesStride, 64);
    }
    if (idx.empty()) return;
    g.idx.reserve(idx.size());
    g.delta.reserve(mats.size());
    SyncOnUnwind sync(m_stream);
    check(dpc_memcpy_h2d(g.idx.p, idx.data(), idx.size() * 4, m_stream), "upload delta index");
    check(dpc_memcpy_h2d(g.delta.p, mats.data(), mats.size() * 4, m_stream), "upload delta");
    check(dpc_matrices_scatter(g.planes.p, g.m_matricesCount, g.idx.p, g.delta.p, idx.size(), m_stream),
          "scatter delta");
    check(dpc_stream_sync(m_stream), "upload sync");
  }

  void remapToTree(GroupImpl& g)
  {
    const size_t n = g.m_objects.size();
    g.xform_slot.reserve(std::max<size_t>(n, 1));
    // The device slot table covers the tree's layout at its last compute
    // (deviceLd entries); an index added since then maps to no slot, and the
    // object is treated like one whose matrix index lies outside the array.
    if (n)
      check(dpc_gather_u32(g.m_tree->deviceSlotOfIndex(), g.m_tree->deviceLd(), g.xform.p, n, g.xform_slot.p,
                           m_stream),
            "map indices to tree slots");
  }

  // Brings the device arrays up to date before a cull.
  void prepare(GroupImpl& g)
  {
    if (g.m_tree) {
      // matrices come from the tree (already on the device); objects may change
      if (g.m_objectsDirty || g.deviceObjects != g.m_objects.size() || g.m_matricesCount != g.m_tree->getTransformCount()) {
        g.m_matricesCount = g.m_tree->getTransformCount();
        uploadObjects(g, g.box_a, g.box_b, g.xform);
        g.deviceObjects = g.m_objects.size();
        g.m_objectsDirty = false;
      }
      remapToTree(g);
      return;
    }
    if (!m_options.refreshOnObjectChange) {
      // The CPU backend's cache (GroupCPU::updateOBBs, ManagerImpl.cpp:114-162):
      // everything is re-read from the objects and the caller's matrices when
      // the cache is dirty, and reused unchanged otherwise.
      g.m_obbDirty |= (g.m_incarnationOBB != g.m_incarnation);
      if (g.m_obbDirty) {
        uploadObjects(g, g.box_a, g.box_b, g.xform);
        uploadAllMatrices(g, g.planes);
        g.deviceObjects = g.m_objects.size();
        g.m_incarnationOBB = g.m_incarnation;
        g.m_obbDirty = false;
        g.m_matricesChanged = false;
        std::fill(g.m_dirtyMatrices.begin(), g.m_dirtyMatrices.end(), 0);
        g.m_objectsDirty = false;
      }
      return;
    }
    // Upload protocol of the reference's GPU sibling (opengl/src/GroupImpl.cpp:68-183):
    // objects when they changed, all matrices when the array changed, otherwise
    // only the matrices marked by groupMatrixChanged.
    if (g.m_objectsDirty || g.deviceObjects != g.m_objects.size()) {
      uploadObjects(g, g.box_a, g.box_b, g.xform);
      g.deviceObjects = g.m_objects.size();
      g.m_objectsDirty = false;
    }
    if (g.m_matricesChanged) {
      uploadAllMatrices(g, g.planes);
      g.m_matricesChanged = false;
    } else {
      uploadDirtyMatrices(g);
    }
    std::fill(g.m_dirtyMatrices.begin(), g.m_dirtyMatrices.end(), 0);
    g.m_obbDirty = false;
  }

  // ResultBitSet::updateChanged (ResultBitSet.cpp:61-108) with the diff on the device.
  void updateResult(GroupImpl& g, ResultImpl& r, const uint64_t* d_new, size_t n)
  {
    updateResult(g, r, d_new, n, [&](ResultImpl& res) {
      check(dpc_result_diff(res.d_old.p, d_new, n, res.d_changed.p, res.d_count.p, 1, res.d_ws.p, m_stream),
            "dpc_result_diff");
    });
  }

  // The same with the step that leaves the new mask in d_new and the diff's
  // outputs in the result's buffers given by the caller (diff).
  template <class Diff>
  void updateResult(GroupImpl& g, ResultImpl& r, const uint64_t* d_new, size_t n, Diff&& diff)
  {
    TraceRange trace("ResultBitSet::updateChanged");
    if (r.m_incarnation != g.m_incarnation) {
      r.m_incarnation = g.m_incarnation;
      const size_t old = r.m_size;
      r.resize(n);
      for (size_t i = old; i < n; ++i) r.setBit(i, true);  // new objects start visible
      r.m_deviceStale = true;
    }
    const size_t words = (n + 63) / 64;
    r.d_old.reserve(std::max<size_t>(words, 1));
    r.d_changed.reserve(std::max<size_t>(n, 1));
    r.d_count.reserve(1);
    const size_t ws = dpc_cull_diff_workspace(n);  // (at least dpc_result_diff_workspace(n))
    r.d_ws.reserve(std::max<size_t>(ws, 16));
    // The result changes only when every step succeeded: after a failure the
    // host bits are still the last successful cull's, and the device copy,
    // which the diff may already have overwritten, is uploaded again next time.
    std::vector<uint64_t> bits(words);
    std::vector<uint32_t> idx;
    try {
      if (r.m_deviceStale && words) {
        check(dpc_memcpy_h2d(r.d_old.p, r.m_bits.data(), words * 8, m_stream), "upload result bits");
        r.m_deviceStale = false;
      }
      diff(r);
      uint32_t count = 0;
      check(dpc_memcpy_d2h(&count, r.d_count.p, 4, m_stream), "download count");
      if (words) check(dpc_memcpy_d2h(bits.data(), d_new, words * 8, m_stream), "download result bits");
      check(dpc_stream_sync(m_stream), "result sync");
      idx.resize(count);
      if (count) {
        check(dpc_memcpy_d2h(idx.data(), r.d_changed.p, count * 4, m_stream), "download changed");
        check(dpc_stream_sync(m_stream), "result sync");
      }
    } catch (...) {
      (void)dpc_stream_sync(m_stream);  // the host buffers above must outlive the copies already queued
      r.m_deviceStale = true;
      throw;
    }
    std::copy(bits.begin(), bits.end(), r.m_bits.begin());
    r.m_changed.clear();
    r.m_changed.reserve(idx.size());
    for (uint32_t i : idx) r.m_changed.push_back(g.m_objects[i]);
  }

  void computeBoundingBox(GroupImpl& g)
  {
    const size_t n = g.m_objects.size();
    DevBuf<float> a, b, planes;
    DevBuf<uint32_t> x;
    const float *pa, *pb, *pp;
    const uint32_t* px;
    size_t ld;
    if (g.m_tree) {
      prepare(g);
      pa = g.box_a.p;
      pb = g.box_b.p;
      px = g.xform_slot.p;
      pp = g.m_tree->deviceWorldPlanes();
      ld = g.m_tree->deviceLd();
    } else if (m_options.refreshOnObjectChange) {
      prepare(g);
      pa = g.box_a.p;
      pb = g.box_b.p;
      px = g.xform.p;
      pp = g.planes.p;
      ld = g.m_matricesCount;
    } else {
      // ManagerBitSet::calculateBoundingBox reads the objects and the caller's
      // matrices as they are now, not the OBB cache (ManagerBitSet.cpp:268-306).
      uploadObjects(g, a, b, x);
      uploadAllMatrices(g, planes);
      pa = a.p;
      pb = b.p;
      px = x.p;
      pp = planes.p;
      ld = g.m_matricesCount;
    }
    g.bbox_ws.reserve(std::max<size_t>(dpc_bbox_workspace(n), 4));
    g.bbox_out.reserve(6);
    SyncOnUnwind sync(m_stream);  // (the launch reads a, b, x and planes, freed on the way out)
    check(dpc_bbox(pa, pb, px, n, pp, ld, g.bbox_out.p, g.bbox_ws.p, m_stream), "dpc_bbox");
    float out[6];
    check(dpc_memcpy_d2h(out, g.bbox_out.p, sizeof out, m_stream), "download bbox");
    check(dpc_stream_sync(m_stream), "bbox sync");
    for (int c = 0; c < 3; ++c) {
      g.m_boundingBox.lower[c] = out[c];
      g.m_boundingBox.upper[c] = out[3 + c];
    }
  }

  Options m_options;
  void* m_stream = nullptr;
  // batched launches
  DevBuf<float> m_vps;
  DevBuf<uint64_t> m_batchMasks;
};

void ManagerImpl::cullBatch(std::vector<CullRequest> const& requests)
{
  TraceRange trace("cullBatch");
  if (requests.empty()) return;
  // one scene per distinct group; one pair per request
  std::vector<std::shared_ptr<GroupImpl>> groups;
  std::vector<uint32_t> group_of(requests.size());
  for (size_t k = 0; k < requests.size(); ++k) {
    auto g = impl(requests[k].group);
    auto r = impl(requests[k].result);
    if (r->m_group != g) throw std::invalid_argument("cullBatch: a result belongs to another group");
    auto it = std::find(groups.begin(), groups.end(), g);
    group_of[k] = static_cast<uint32_t>(it - groups.begin());
    if (it == groups.end()) groups.push_back(g);
  }
  std::vector<dpc_scene_desc> scenes(groups.size());
  std::vector<dpc_group_range> ranges(groups.size());
  for (size_t s = 0; s < groups.size(); ++s) {
    GroupImpl& g = *groups[s];
    prepare(g);
    const size_t n = g.m_objects.size();
    scenes[s].box_a = g.box_a.p;
    scenes[s].box_b = g.box_b.p;
    scenes[s].xform = g.m_tree ? g.xform_slot.p : g.xform.p;
    scenes[s].planes = g.m_tree ? g.m_tree->deviceWorldPlanes() : g.planes.p;
    scenes[s].ld = g.m_tree ? g.m_tree->deviceLd() : g.m_matricesCount;
    scenes[s].n_objects = n;
    ranges[s] = dpc_group_range{static_cast<uint32_t>(s), 0, 0, n};
  }
  std::vector<dpc_cull_pair> pairs(requests.size());
  std::vector<uint64_t> offset(requests.size());
  uint64_t words = 0;
  std::vector<float> vps(16 * requests.size());
  for (size_t k = 0; k < requests.size(); ++k) {
    offset[k] = words;
    words += (groups[group_of[k]]->m_objects.size() + 63) / 64;
    pairs[k] = dpc_cull_pair{static_cast<uint32_t>(k), group_of[k], offset[k]};
    std::memcpy(&vps[16 * k], requests[k].viewProjection.m, 64);
  }
  dpc_batch_plan* plan = nullptr;
  check(dpc_batch_plan_create(ranges.data(), ranges.size(), pairs.data(), pairs.size(),
                              static_cast<uint32_t>(requests.size()), nullptr, &plan),
        "dpc_batch_plan_create");
  struct PlanGuard {
    dpc_batch_plan* p;
    ~PlanGuard() { dpc_batch_plan_destroy(p); }
  } guard{plan};
  check(dpc_batch_plan_bind(plan, scenes.data(), scenes.size()), "dpc_batch_plan_bind");
  m_vps.reserve(vps.size());
  m_batchMasks.reserve(std::max<uint64_t>(words, 1));
  SyncOnUnwind sync(m_stream);
  check(dpc_memcpy_h2d(m_vps.p, vps.data(), vps.size() * 4, m_stream), "upload frusta");
  check(dpc_cull_batch(plan, m_vps.p, m_batchMasks.p, std::max<uint64_t>(words, 1), m_stream), "dpc_cull_batch");
  // results in request order, exactly as the same sequence of cull() calls
  for (size_t k = 0; k < requests.size(); ++k) {
    GroupImpl& g = *groups[group_of[k]];
    updateResult(g, *impl(requests[k].result), m_batchMasks.p + offset[k], g.m_objects.size());
  }
}

}  // namespace

culling::Manager* Manager::create(Options const& options) { return new ManagerImpl(options); }

void groupSetTreeMatrices(culling::Manager* manager, GroupSharedPtr const& group, transform::Tree const& tree)
{
  auto* m = dynamic_cast<ManagerImpl*>(manager);
  if (!m) throw std::invalid_argument("groupSetTreeMatrices: not a HIP culling manager");
  m->bindTree(group, tree);
}

void cullBatch(culling::Manager* manager, std::vector<CullRequest> const& requests)
{
  auto* m = dynamic_cast<ManagerImpl*>(manager);
  if (!m) throw std::invalid_argument("cullBatch: not a HIP culling manager");
  m->cullBatch(requests);
}

}  // namespace hip
}  // namespace culling
}  // namespace dpcull

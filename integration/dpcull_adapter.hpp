// dpcull_adapter.hpp -- the HIP culling backend behind nvpro-pipeline's own
// culling interface, dp::culling::Manager (dp/culling/Manager.h:108-143), and the
// transform-tree hookup that xbar's culling declares but never attaches
// (CullingImpl::TransformObserver, dp/sg/xbar/culling/inc/CullingImpl.h:99-111;
// its forwarding body is CullingImpl.cpp:216-223).
//
// ManagerAdapter<Front> implements Front::Manager -- every one of its eighteen
// methods -- by forwarding to a dpcull::culling::Manager created by
// dpcull::culling::hip::Manager::create() (include/dpcull/Manager.hpp).  Front
// is a traits type that names the front end's classes and converts its value
// types:
//
//   struct Front {
//     using Manager = ...;  using Object = ...;  using Group = ...;
//     using Result = ...;   using Payload = ...; using Box3f = ...; using Mat44f = ...;
//     static dpcull::Box3f  toBack(Box3f const&);
//     static Box3f          toFront(dpcull::Box3f const&);
//     static dpcull::Mat44f toBack(Mat44f const&);
//   };
//
// dp/culling/hip/src/ManagerImpl.cpp instantiates it with dp::culling's types
// (that file compiles inside the reference tree); tests/cpp/integration_test.cpp
// instantiates it with dpcull::culling's own interface as the front end, which
// has the same method set, and drives it with xbar's per-frame call sequence.
//
// Handles.  Every front-end handle wraps one backend handle.  The backend
// object's payload points back (weakly) at the front-end object, so the
// backend's changed lists and groupGetObject map back to the front-end objects
// the caller created; the caller's own payload stays on the front-end object,
// where objectGetUserData finds it (CullingImpl.cpp:139-143 reads it for every
// changed object).
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <vector>

#include "dpcull/Manager.hpp"
#include "dpcull/Tree.hpp"

namespace dpcull {
namespace adapter {

template <typename Front>
class ManagerAdapter : public Front::Manager {
 public:
  using FObjectSharedPtr = std::shared_ptr<typename Front::Object>;
  using FGroupSharedPtr = std::shared_ptr<typename Front::Group>;
  using FResultSharedPtr = std::shared_ptr<typename Front::Result>;
  using FPayloadSharedPtr = std::shared_ptr<typename Front::Payload>;

  explicit ManagerAdapter(culling::hip::Options const& options = culling::hip::Options())
      : m_back(culling::hip::Manager::create(options))
  {
  }

  // ---- objects (ManagerBitSet.cpp:49-109) ----------------------------------
  FObjectSharedPtr objectCreate(FPayloadSharedPtr const& userData) override
  {
    auto front = std::make_shared<ObjectH>();
    auto back = std::make_shared<BackPayload>();
    back->front = front;
    front->back = m_back->objectCreate(back);
    front->user = userData;
    return front;
  }

  void objectSetBoundingBox(FObjectSharedPtr const& object, typename Front::Box3f const& boundingBox) override
  {
    m_back->objectSetBoundingBox(obj(object)->back, Front::toBack(boundingBox));
  }

  void objectSetTransformIndex(FObjectSharedPtr const& object, size_t index) override
  {
    m_back->objectSetTransformIndex(obj(object)->back, index);
  }

  void objectSetUserData(FObjectSharedPtr const& object, FPayloadSharedPtr const& userData) override
  {
    obj(object)->user = userData;
  }

  FPayloadSharedPtr const& objectGetUserData(FObjectSharedPtr const& object) override { return obj(object)->user; }

  // ---- groups (ManagerBitSet.cpp:111-149, GroupBitSet.cpp:76-135) -----------
  FGroupSharedPtr groupCreate() override
  {
    auto front = std::make_shared<GroupH>();
    front->back = m_back->groupCreate();
    return front;
  }

  void groupAddObject(FGroupSharedPtr const& group, FObjectSharedPtr const& object) override
  {
    m_back->groupAddObject(grp(group)->back, obj(object)->back);
  }

  FObjectSharedPtr groupGetObject(FGroupSharedPtr const& group, size_t index) override
  {
    return frontOf(m_back->groupGetObject(grp(group)->back, index));
  }

  void groupRemoveObject(FGroupSharedPtr const& group, FObjectSharedPtr const& object) override
  {
    m_back->groupRemoveObject(grp(group)->back, obj(object)->back);
  }

  size_t groupGetCount(FGroupSharedPtr const& group) override { return m_back->groupGetCount(grp(group)->back); }

  void groupSetMatrices(FGroupSharedPtr const& group, void const* matrices, size_t numberOfMatrices,
                        size_t stride) override
  {
    m_back->groupSetMatrices(grp(group)->back, matrices, numberOfMatrices, stride);
  }

  void groupMatrixChanged(FGroupSharedPtr const& group, size_t index) override
  {
    m_back->groupMatrixChanged(grp(group)->back, index);
  }

  FResultSharedPtr groupCreateResult(FGroupSharedPtr const& group) override
  {
    auto front = std::make_shared<ResultH>();
    front->back = m_back->groupCreateResult(grp(group)->back);
    return front;
  }

  // ---- results (ResultBitSet.cpp:61-128, ResultBitSet.h:69-76) --------------
  std::vector<FObjectSharedPtr> const& resultGetChanged(FResultSharedPtr const& result) override
  {
    return res(result)->changed;
  }

  bool resultObjectIsVisible(FResultSharedPtr const& result, FObjectSharedPtr const& object) override
  {
    return m_back->resultObjectIsVisible(res(result)->back, obj(object)->back);
  }

  // ---- cull and bounding box (cpu/src/ManagerImpl.cpp:467-518, ManagerBitSet.cpp:151-162)
  void cull(FGroupSharedPtr const& group, FResultSharedPtr const& result,
            typename Front::Mat44f const& viewProjection) override
  {
    ResultH* r = res(result);
    m_back->cull(grp(group)->back, r->back, Front::toBack(viewProjection));
    // the changed list in the backend's order (ascending object index), mapped back
    auto const& changed = m_back->resultGetChanged(r->back);
    r->changed.clear();
    r->changed.reserve(changed.size());
    for (auto const& o : changed) r->changed.push_back(frontOf(o));
  }

  typename Front::Box3f getBoundingBox(FGroupSharedPtr const& group) const override
  {
    return Front::toFront(m_back->getBoundingBox(grp(group)->back));
  }

  // ---- beyond the reference interface ---------------------------------------
  // The group reads the world matrices of a GPU transform tree where they are
  // computed (no host copy); object transform indices are tree indices.
  void groupSetTreeMatrices(FGroupSharedPtr const& group, transform::Tree const& tree)
  {
    culling::hip::groupSetTreeMatrices(m_back.get(), grp(group)->back, tree);
  }

  culling::Manager& backend() { return *m_back; }

 private:
  struct ObjectH;
  struct BackPayload : culling::Payload {
    std::weak_ptr<ObjectH> front;
  };
  struct ObjectH : Front::Object {
    culling::ObjectSharedPtr back;
    FPayloadSharedPtr user;
  };
  struct GroupH : Front::Group {
    culling::GroupSharedPtr back;
  };
  struct ResultH : Front::Result {
    culling::ResultSharedPtr back;
    std::vector<FObjectSharedPtr> changed;
  };

  // Handles from another manager are a caller error the reference does not
  // check either (it static_casts); these casts assume this manager's handles.
  static ObjectH* obj(FObjectSharedPtr const& o)
  {
    if (!o) throw std::invalid_argument("dp::culling::hip: null object");
    return static_cast<ObjectH*>(o.get());
  }
  static GroupH* grp(FGroupSharedPtr const& g)
  {
    if (!g) throw std::invalid_argument("dp::culling::hip: null group");
    return static_cast<GroupH*>(g.get());
  }
  static ResultH* res(FResultSharedPtr const& r)
  {
    if (!r) throw std::invalid_argument("dp::culling::hip: null result");
    return static_cast<ResultH*>(r.get());
  }
  FObjectSharedPtr frontOf(culling::ObjectSharedPtr const& back) const
  {
    if (!back) return FObjectSharedPtr();
    auto const& p = m_back->objectGetUserData(back);
    return std::static_pointer_cast<BackPayload>(p)->front.lock();
  }

  std::unique_ptr<culling::Manager> m_back;
};

// CullingImpl::TransformObserver (CullingImpl.h:99-111, CullingImpl.cpp:216-223):
// every world matrix a Tree::compute() changed is reported to the culling
// group with groupMatrixChanged, so a group whose matrix array keeps its
// address, count and stride (groupSetMatrices then changes nothing,
// GroupBitSet.cpp:121-135) re-reads exactly the matrices that moved.  The
// reference declares this observer but never attaches it; attach this one to
// the tree right after the group is created.
template <typename ManagerT, typename GroupPtrT>
class TransformForwarder : public transform::TreeObserver {
 public:
  TransformForwarder(ManagerT& manager, GroupPtrT group) : m_manager(manager), m_group(std::move(group)) {}

  void onWorldMatricesChanged(const std::vector<uint64_t>& bits, size_t count) override
  {
    // BitArray::traverseBits order: ascending indices
    for (size_t w = 0; w < bits.size(); ++w) {
      uint64_t x = bits[w];
      while (x) {
        const size_t i = w * 64 + static_cast<size_t>(__builtin_ctzll(x));
        if (i < count) m_manager.groupMatrixChanged(m_group, i);
        x &= x - 1;
      }
    }
  }

 private:
  ManagerT& m_manager;
  GroupPtrT m_group;
};

}  // namespace adapter
}  // namespace dpcull

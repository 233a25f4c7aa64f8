// cpu_manager.hpp -- CPU restatement of nvpro-pipeline's culling manager
// (dp::culling::cpu::ManagerImpl with ManagerBitSet, GroupBitSet/GroupCPU,
// ObjectBitSet and ResultBitSet).
//
// TEST INFRASTRUCTURE ONLY: the checker that tests/cpp compares the HIP manager
// against.  It implements the same abstract interface (include/dpcull/Manager.hpp)
// so one trace of calls can drive both.  All arithmetic goes through oracle.c.
//
// Stateful behaviour restated (reference file:line):
//   objectSetBoundingBox/TransformIndex store the values; they dirty nothing,
//     because ObjectBitSet::getGroup() is never set (ManagerBitSet.cpp:88-109,
//     GroupBitSet.cpp:76-91) -- unless refreshOnObjectChange is set, which marks
//     the group's OBB cache and bounding box dirty (the HIP backend's default);
//   addObject / removeObject (GroupBitSet.cpp:76-119): append / swap-remove, bump
//     the incarnation, dirty the OBB cache and bounding box, notify results;
//   setMatrices (GroupBitSet.cpp:121-135): on a new pointer, count or stride,
//     dirty everything; markMatrixDirty (GroupBitSet.h:140-150) for index < count;
//   GroupCPU::updateOBBs (cpu/src/ManagerImpl.cpp:114-162): recompute every OBB
//     from the current objects and current matrix memory when dirty;
//   cull (ManagerImpl.cpp:467-518) then ResultBitSet::updateChanged
//     (ResultBitSet.cpp:61-108); isVisible (ResultBitSet.h:69-76);
//   getBoundingBox (ManagerBitSet.cpp:151-162) cached until dirty, computed
//     from the current objects and matrices (ManagerBitSet.cpp:268-306).
//
// Parity status: the reference's culling manager does not compile in this image
// (it includes Boost headers), so this restatement's stateful behaviour is
// pinned by the hand-traced scenarios in tests/cpp/manager_parity.cpp; its
// arithmetic is oracle.c's.
#pragma once

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../include/dpcull/Manager.hpp"
#include "oracle.h"

namespace oracle {

using namespace dpcull;
using namespace dpcull::culling;

class CpuManager : public culling::Manager {
 public:
  explicit CpuManager(bool refreshOnObjectChange = false) : m_refresh(refreshOnObjectChange) {}

  struct Grp;
  struct Obj : Object {
    float lower[4] = {0, 0, 0, 1};
    float extent[4] = {0, 0, 0, 0};
    size_t transformIndex = ~size_t(0);
    PayloadSharedPtr userData;
    size_t groupIndex = ~size_t(0);
    std::weak_ptr<Object> self;
    Grp* group = nullptr;  // only used when refreshOnObjectChange
  };
  struct Res;
  struct Grp : Group {
    std::vector<std::shared_ptr<Obj>> objects;
    const void* matrices = nullptr;
    size_t count = 0, stride = 0;
    bool obbDirty = true, bboxDirty = true;
    size_t incarnation = 0, incarnationOBB = ~size_t(0);
    std::vector<float> obbs;
    Box3f bbox;
    std::vector<Res*> results;
    ~Grp() override
    {
      for (auto& o : objects) {
        o->groupIndex = ~size_t(0);
        o->group = nullptr;
      }
    }
  };
  struct Res : Result {
    std::shared_ptr<Grp> group;
    std::vector<ObjectSharedPtr> changed;
    size_t incarnation = ~size_t(0);
    std::vector<uint8_t> bits;  // one byte per bit; BitArray semantics
    ~Res() override
    {
      auto& r = group->results;
      r.erase(std::remove(r.begin(), r.end(), this), r.end());
    }
  };

  ObjectSharedPtr objectCreate(PayloadSharedPtr const& userData) override
  {
    auto o = std::make_shared<Obj>();
    o->userData = userData;
    return o;
  }
  void objectSetBoundingBox(ObjectSharedPtr const& object, Box3f const& b) override
  {
    auto* o = obj(object);
    for (int c = 0; c < 3; ++c) {
      o->lower[c] = b.lower[c];
      o->extent[c] = b.upper[c] - b.lower[c];
    }
    o->lower[3] = 1.0f;
    o->extent[3] = 0.0f;
    if (m_refresh && o->group) o->group->obbDirty = o->group->bboxDirty = true;
  }
  void objectSetTransformIndex(ObjectSharedPtr const& object, size_t index) override
  {
    auto* o = obj(object);
    o->transformIndex = index;
    if (m_refresh && o->group) o->group->obbDirty = o->group->bboxDirty = true;
  }
  void objectSetUserData(ObjectSharedPtr const& object, PayloadSharedPtr const& p) override { obj(object)->userData = p; }
  PayloadSharedPtr const& objectGetUserData(ObjectSharedPtr const& object) override { return obj(object)->userData; }

  GroupSharedPtr groupCreate() override { return std::make_shared<Grp>(); }
  void groupAddObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) override
  {
    auto g = grp(group);
    auto o = std::static_pointer_cast<Obj>(object);
    if (o->groupIndex != ~size_t(0)) throw std::runtime_error("object already belongs to a different group");
    o->groupIndex = g->objects.size();
    o->group = g.get();
    g->objects.push_back(o);
    ++g->incarnation;
    g->bboxDirty = g->obbDirty = true;
  }
  ObjectSharedPtr groupGetObject(GroupSharedPtr const& group, size_t index) override
  {
    return grp(group)->objects.at(index);
  }
  void groupRemoveObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) override
  {
    auto g = grp(group);
    auto o = std::static_pointer_cast<Obj>(object);
    const size_t old = o->groupIndex;
    if (!(old < g->objects.size() && g->objects[old] == o)) throw std::runtime_error("object does not belong to this group");
    g->objects[old] = g->objects.back();
    g->objects[old]->groupIndex = old;
    o->groupIndex = ~size_t(0);
    o->group = nullptr;
    g->objects.pop_back();
    const size_t moved_from = g->objects.size();
    for (Res* r : g->results) {  // ResultBitSet::onNotify
      if (old < r->bits.size()) r->bits[old] = moved_from < r->bits.size() ? r->bits[moved_from] : 1;
    }
    g->bboxDirty = g->obbDirty = true;
    ++g->incarnation;
  }
  size_t groupGetCount(GroupSharedPtr const& group) override { return grp(group)->objects.size(); }
  void groupSetMatrices(GroupSharedPtr const& group, void const* m, size_t count, size_t stride) override
  {
    auto g = grp(group);
    if (m != g->matrices || count != g->count || stride != g->stride) {
      g->matrices = m;
      g->count = count;
      g->stride = stride;
      g->bboxDirty = g->obbDirty = true;
    }
  }
  void groupMatrixChanged(GroupSharedPtr const& group, size_t index) override
  {
    auto g = grp(group);
    if (index < g->count) g->bboxDirty = g->obbDirty = true;
  }
  ResultSharedPtr groupCreateResult(GroupSharedPtr const& group) override
  {
    auto r = std::make_shared<Res>();
    r->group = grp(group);
    r->group->results.push_back(r.get());
    return r;
  }
  std::vector<ObjectSharedPtr> const& resultGetChanged(ResultSharedPtr const& result) override
  {
    return res(result)->changed;
  }
  bool resultObjectIsVisible(ResultSharedPtr const& result, ObjectSharedPtr const& object) override
  {
    auto r = res(result);
    const size_t i = obj(object)->groupIndex;
    return i < r->bits.size() ? r->bits[i] != 0 : true;
  }

  void cull(GroupSharedPtr const& group, ResultSharedPtr const& result, Mat44f const& vp) override
  {
    auto g = grp(group);
    auto r = res(result);
    updateOBBs(*g);
    const size_t n = g->objects.size();
    std::vector<uint8_t> visible(n);
    for (size_t i = 0; i < n; ++i) visible[i] = static_cast<uint8_t>(orc_is_visible_obb(vp.m, &g->obbs[16 * i]));
    // ResultBitSet::updateChanged
    if (r->incarnation != g->incarnation) {
      r->incarnation = g->incarnation;
      const size_t old = r->bits.size();
      r->bits.resize(n, 0);
      for (size_t i = old; i < n; ++i) r->bits[i] = 1;
    }
    r->changed.clear();
    for (size_t i = 0; i < n; ++i)
      if (visible[i] != r->bits[i]) r->changed.push_back(g->objects[i]);
    r->bits = visible;
  }

  Box3f getBoundingBox(GroupSharedPtr const& group) const override
  {
    auto g = std::static_pointer_cast<Grp>(group);
    if (g->bboxDirty) {
      const size_t n = g->objects.size();
      std::vector<float> l4(4 * n), e4(4 * n);
      std::vector<uint64_t> x(n);
      for (size_t i = 0; i < n; ++i) {
        std::memcpy(&l4[4 * i], g->objects[i]->lower, 16);
        std::memcpy(&e4[4 * i], g->objects[i]->extent, 16);
        x[i] = g->objects[i]->transformIndex;
      }
      float out[6];
      orc_bbox_group(n, l4.data(), e4.data(), x.data(), g->matrices, g->stride, out);
      for (int c = 0; c < 3; ++c) {
        g->bbox.lower[c] = out[c];
        g->bbox.upper[c] = out[3 + c];
      }
      g->bboxDirty = false;
    }
    return g->bbox;
  }

 private:
  Obj* obj(ObjectSharedPtr const& o) { return static_cast<Obj*>(o.get()); }
  std::shared_ptr<Grp> grp(GroupSharedPtr const& g) { return std::static_pointer_cast<Grp>(g); }
  std::shared_ptr<Res> res(ResultSharedPtr const& r) { return std::static_pointer_cast<Res>(r); }

  void updateOBBs(Grp& g)
  {
    g.obbDirty |= (g.incarnationOBB != g.incarnation);
    if (!g.obbDirty) return;
    const size_t n = g.objects.size();
    g.obbs.resize(16 * n);
    for (size_t i = 0; i < n; ++i) {
      const Obj& o = *g.objects[i];
      const float* m = reinterpret_cast<const float*>(static_cast<const char*>(g.matrices) + o.transformIndex * g.stride);
      orc_obb_build(o.lower, o.extent, m, &g.obbs[16 * i]);
    }
    g.incarnationOBB = g.incarnation;
    g.obbDirty = false;
  }

  bool m_refresh;
};

}  // namespace oracle

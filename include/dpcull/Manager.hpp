// Manager.hpp -- C++ mirror of dp::culling::Manager (dp/culling/Manager.h:38-143)
// with a HIP backend, dpcull::culling::hip::Manager::create(), the sibling of the
// reference's cpu::Manager::create() (dp/culling/cpu/Manager.h:39-43) and
// opengl::Manager::create() (dp/culling/opengl/Manager.h:38-42).
//
// Handles are shared pointers to opaque objects, groups and results, exactly as
// in the reference.  Behaviour follows ManagerBitSet / GroupBitSet / ObjectBitSet /
// ResultBitSet (dp/culling/src/*.cpp):
//   - an object belongs to at most one group; adding it twice throws
//     std::runtime_error("object already belongs to a different group");
//     removing an object from a group it is not in throws (GroupBitSet.cpp:76-119);
//   - removal swap-removes (the last object takes the freed position) and results
//     move that object's visibility bit along (ResultBitSet.cpp:110-128);
//   - a result's objects are visible until culled; after cull(), resultGetChanged
//     lists the objects whose visibility flipped, in index order
//     (ResultBitSet.cpp:61-108); resultObjectIsVisible is true for objects the
//     result has not culled yet (ResultBitSet.h:69-76);
//   - matrices are borrowed from the caller: groupSetMatrices(ptr, count, stride)
//     re-reads them all when any of the three changes, groupMatrixChanged(i)
//     re-reads matrix i (GroupBitSet.cpp:121-135, GroupBitSet.h:140-150);
//   - getBoundingBox is cached until the group changes (ManagerBitSet.cpp:151-162).
//
// Box and transform-index updates of an object that is already in a group.
// The reference never links an object to its group (GroupBitSet::addObject does
// not call ObjectBitSet::setGroup, GroupBitSet.cpp:76-91), so such updates do not
// dirty the CPU backend's cached OBBs: the CPU result keeps using the old box
// until something else (an add/remove, new matrices, a changed matrix) refreshes
// the whole cache.  This backend applies such updates at the next cull by default
// (Options::refreshOnObjectChange = true).  With refreshOnObjectChange = false it
// reproduces the CPU backend's caching exactly, for bit-for-bit comparisons
// against it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "Math.hpp"

namespace dpcull {

namespace transform {
class Tree;
}

namespace culling {

class Object {
 public:
  virtual ~Object() = default;
};
class Group {
 public:
  virtual ~Group() = default;
};
class Result {
 public:
  virtual ~Result() = default;
};
class Payload {
 public:
  virtual ~Payload() = default;
};

using ObjectSharedPtr = std::shared_ptr<Object>;
using GroupSharedPtr = std::shared_ptr<Group>;
using ResultSharedPtr = std::shared_ptr<Result>;
using PayloadSharedPtr = std::shared_ptr<Payload>;

// dp::culling::Mode (dp/culling/Manager.h:94-100) plus the HIP backend.
enum class Mode { CPU, OPENGL_COMPUTE, CUDA, AUTO, HIP };

class Manager {
 public:
  virtual ~Manager() = default;

  virtual ObjectSharedPtr objectCreate(PayloadSharedPtr const& userData) = 0;
  virtual void objectSetBoundingBox(ObjectSharedPtr const& object, Box3f const& boundingBox) = 0;
  virtual void objectSetTransformIndex(ObjectSharedPtr const& object, size_t index) = 0;
  virtual void objectSetUserData(ObjectSharedPtr const& object, PayloadSharedPtr const& userData) = 0;
  virtual PayloadSharedPtr const& objectGetUserData(ObjectSharedPtr const& object) = 0;

  virtual GroupSharedPtr groupCreate() = 0;
  virtual void groupAddObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) = 0;
  virtual ObjectSharedPtr groupGetObject(GroupSharedPtr const& group, size_t index) = 0;
  virtual void groupRemoveObject(GroupSharedPtr const& group, ObjectSharedPtr const& object) = 0;
  virtual size_t groupGetCount(GroupSharedPtr const& group) = 0;
  virtual void groupSetMatrices(GroupSharedPtr const& group, void const* matrices, size_t numberOfMatrices,
                                size_t stride) = 0;
  virtual void groupMatrixChanged(GroupSharedPtr const& group, size_t index) = 0;
  virtual ResultSharedPtr groupCreateResult(GroupSharedPtr const& group) = 0;

  virtual std::vector<ObjectSharedPtr> const& resultGetChanged(ResultSharedPtr const& result) = 0;
  virtual bool resultObjectIsVisible(ResultSharedPtr const& result, ObjectSharedPtr const& object) = 0;

  virtual void cull(GroupSharedPtr const& group, ResultSharedPtr const& result, Mat44f const& viewProjection) = 0;
  virtual Box3f getBoundingBox(GroupSharedPtr const& group) const = 0;
};

namespace hip {

struct Options {
  // Apply objectSetBoundingBox / objectSetTransformIndex to objects already in a
  // group at the next cull (true), or cache like the CPU backend (false).
  bool refreshOnObjectChange = true;
};

class Manager {
 public:
  // A culling manager on the current HIP device.
  static culling::Manager* create(Options const& options = Options());
};

// Extensions with no reference counterpart.
//
// Bind a group's matrices to a GPU transform tree: the group reads the tree's
// world matrices where they are computed, with no host round trip.  Object
// transform indices are tree indices.  Call after every Tree::compute() that
// changed the tree's structure (it rebinds the slot map).
void groupSetTreeMatrices(culling::Manager* manager, GroupSharedPtr const& group, transform::Tree const& tree);

// Cull many (group, result, viewProjection) triples in one batched launch.
// Results are updated exactly as by the same sequence of cull() calls.
struct CullRequest {
  GroupSharedPtr group;
  ResultSharedPtr result;
  Mat44f viewProjection;
};
void cullBatch(culling::Manager* manager, std::vector<CullRequest> const& requests);

}  // namespace hip
}  // namespace culling
}  // namespace dpcull

// Tree.hpp -- C++ mirror of dp::transform::Tree (dp/transform/Tree.h:42-129)
// whose world matrices live on the GPU and are propagated by the
// level-synchronous kernels behind dpc_tree_propagate.
//
// Same surface and semantics as the reference:
//   - index 0 is the virtual root (identity), getRoot() == 0;
//   - addTransform takes the lowest free index (Tree.cpp:100-110), growing the
//     arrays by 65536 entries at a time (Tree.cpp:35,117-126), sets
//     level = level[parent] + 1 and appends (parent, index) to that level's
//     list (Tree.cpp:54-76); an unknown parent throws std::runtime_error;
//   - removeTransform swap-removes the entry from its level list and frees the
//     index; children are left in place (orphaned) and keep reading the freed
//     index's matrix (Tree.cpp:78-98);
//   - compute() recomputes, level by level in list order, every entry whose
//     local matrix or parent world matrix changed, then reports the changed
//     world matrices to observers and clears the dirty state (Tree.cpp:133-166).
//
// Device layout: the level lists are compiled into "slots" (slot 0 = root, then
// each level's entries, then every index that is in no list), so each level is
// a contiguous slot range and the kernels read and write whole cache lines.  The
// compile runs only after structural changes; a level in which an entry reads a
// matrix that another entry of the same level writes (possible only after index
// reuse) is split into passes in list order, which keeps the sequential order
// of the reference.  Inside a pass no entry reads what another writes, so the
// pass is ordered by parent slot: the children of a slot range are then
// contiguous, which the propagation plan's tile runs use.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "Math.hpp"

namespace dpcull {
namespace transform {

typedef uint32_t Index;

// Receives the set of world matrices changed by one compute()
// (Tree::EventWorldMatricesChanged, dp/transform/Tree.h:48-60).
class TreeObserver {
 public:
  virtual ~TreeObserver() = default;
  // bits: bit i of 64-bit word i/64 is set when world matrix i changed; count = getTransformCount().
  virtual void onWorldMatricesChanged(const std::vector<uint64_t>& bits, size_t count) = 0;
};

struct LevelEntry {
  uint32_t parent;
  uint32_t transform;
};

// Host bookkeeping of the tree (index allocation, level lists) and its
// compiled slot layout.  No device work; usable and testable on its own.
class TreeLayout {
 public:
  TreeLayout();
  Index addTransform(Index parentIndex);  // throws std::runtime_error on an invalid parent
  void removeTransform(Index index);      // throws std::runtime_error on an invalid index
  bool isValidIndex(Index index) const;
  size_t capacity() const { return m_level.size(); }
  const std::vector<std::vector<LevelEntry>>& levels() const { return m_levels; }
  uint32_t level(Index i) const { return m_level[i]; }

  // The compiled layout.
  struct Compiled {
    std::vector<uint32_t> index_of_slot;   // capacity entries
    std::vector<uint32_t> slot_of_index;   // capacity entries
    std::vector<uint32_t> parent_slot;     // capacity entries (0 for slots outside the levels)
    std::vector<uint64_t> level_offsets;   // pass p = slots [off[p], off[p+1]); passes run in order
  };
  Compiled compile() const;

 private:
  Index allocateIndex();
  void grow(size_t newSize);

  std::vector<uint64_t> m_free;  // bit i set = index i free
  std::vector<uint32_t> m_level;
  std::vector<std::vector<LevelEntry>> m_levels;
};

class Tree {
 public:
  // Allocates device memory on the current HIP device.
  Tree();
  ~Tree();
  Tree(const Tree&) = delete;
  Tree& operator=(const Tree&) = delete;

  Index addTransform(Index parentIndex, Mat44f const& matrix);
  void removeTransform(Index transformIndex);
  // The camera argument is unused, as in the reference (billboards are compiled out, Tree.cpp:137-150).
  void compute(Mat44f const& camera = Mat44f::identity());

  // Host copies, indexed by transform index (downloaded after compute on first use).
  Mat44f const& getWorldMatrix(Index index) const;
  Mat44f const* getWorldMatrices() const;
  size_t getTransformCount() const { return m_layout.capacity(); }
  // The bits the last compute reported (the reference clears its bitset after notifying).
  const std::vector<uint64_t>& getLastChangedWorldMatrices() const { return m_lastEvent; }
  void updateLocalMatrix(Index index, Mat44f const& matrix);
  Index getRoot() const { return 0; }

  void attach(TreeObserver* observer);
  void detach(TreeObserver* observer);

  const TreeLayout& layout() const { return m_layout; }

  // Device-side access (no reference counterpart): world matrices as row planes
  // by slot, and the slot of every transform index, for consumers that stay on
  // the GPU (a culling group bound to this tree).
  const float* deviceWorldPlanes() const;
  size_t deviceLd() const;
  const uint32_t* deviceSlotOfIndex() const;
  void* stream() const;
  // World-matrix inverse transposes, dp::math::invertTranspose
  // (dp/math/Matmnt.h:1171-1201), kept current by compute() for every world
  // matrix it changed: the reference renderer computes them on the host, one
  // changed index at a time (ShaderManagerTransformsRiXFx::updateTransforms,
  // dp/sg/renderer/rix/gl/src/ShaderManagerRiXFx.cpp:162-182).  Off by default.
  // A singular world matrix (det == 0) keeps its previous inverse transpose
  // (zero for a new index); the reference leaves its output undefined there.
  void setInverseTransposeEnabled(bool enabled);
  bool inverseTransposeEnabled() const { return m_itEnabled; }
  // Host copy by transform index (downloaded on first use after a compute).
  Mat44f const* getWorldMatricesInverseTranspose() const;
  // Row planes by slot, like deviceWorldPlanes().
  const float* deviceWorldInverseTransposePlanes() const;

  // Number of compute() calls so far; consumers that cache something derived
  // from the world matrices compare it to know when to refresh.
  uint64_t computeCount() const { return m_computeCount; }

 private:
  struct Device;
  void ensureCapacity();
  void recompile();
  void updateInverseTranspose();

  TreeLayout m_layout;
  TreeLayout::Compiled m_compiled;
  bool m_structureDirty = true;
  std::vector<Mat44f> m_local;
  std::vector<uint8_t> m_dirtyLocal;  // by index
  bool m_anyDirty = true;
  mutable std::vector<Mat44f> m_worldHost;
  mutable bool m_worldHostStale = true;
  bool m_itEnabled = false;
  bool m_itAll = true;  // the next compute() recomputes every inverse transpose
  mutable std::vector<Mat44f> m_itHost;
  mutable bool m_itHostStale = true;
  std::vector<uint64_t> m_lastEvent;
  std::vector<TreeObserver*> m_observers;
  std::unique_ptr<Device> m_dev;
  uint64_t m_computeCount = 0;
  bool m_dirtyWorldStale = false;  // a failed compute() may have left dirty-world bytes set
};

}  // namespace transform
}  // namespace dpcull

// Tree.cpp -- host side of the GPU transform tree (see include/dpcull/Tree.hpp).
// Bookkeeping follows dp/transform/src/Tree.cpp; all matrix arithmetic runs on
// the device through the C ABI.
#include "dpcull/Tree.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "dpcull_hip.h"
#include "trace.hpp"
#include "unwind.hpp"

namespace dpcull {
namespace transform {

namespace {

constexpr size_t kGrowth = 65536;  // dp/transform/src/Tree.cpp:35

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
  void swap(DevBuf& o)
  {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
  void reserve(size_t count)
  {
    if (count <= n) return;
    if (p) dpc_free(p);
    p = nullptr;
    n = 0;
    void* q = nullptr;
    check(dpc_malloc(&q, count * sizeof(T)), "dpc_malloc");
    p = static_cast<T*>(q);
    n = count;
  }
  ~DevBuf()
  {
    if (p) dpc_free(p);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// TreeLayout
// ---------------------------------------------------------------------------
TreeLayout::TreeLayout()
{
  grow(kGrowth);
  m_free[0] &= ~uint64_t(1);  // the root is allocated (Tree.cpp:42)
  m_level[0] = 0xFFFFFFFFu;   // Tree.cpp:47; its children land on level 0
}

void TreeLayout::grow(size_t newSize)
{
  const size_t old = m_level.size();
  m_level.resize(newSize, 0);
  m_free.resize((newSize + 63) / 64, 0);
  for (size_t i = old; i < newSize; ++i) m_free[i >> 6] |= uint64_t(1) << (i & 63);  // new entries are free
}

bool TreeLayout::isValidIndex(Index i) const
{
  return i < m_level.size() && !((m_free[i >> 6] >> (i & 63)) & 1);
}

Index TreeLayout::allocateIndex()
{
  // BitArray::countLeadingZeroes: the lowest set (free) bit, or the size (Tree.cpp:100-110).
  size_t idx = m_level.size();
  for (size_t w = 0; w < m_free.size(); ++w) {
    if (m_free[w]) {
      idx = std::min(m_level.size(), w * 64 + static_cast<size_t>(__builtin_ctzll(m_free[w])));
      break;
    }
  }
  if (idx == m_level.size()) grow(m_level.size() + kGrowth);
  m_free[idx >> 6] &= ~(uint64_t(1) << (idx & 63));
  return static_cast<Index>(idx);
}

Index TreeLayout::addTransform(Index parent)
{
  if (!isValidIndex(parent)) throw std::runtime_error("Tree::addTransform: Parent does not exist");
  const Index idx = allocateIndex();
  const uint32_t lvl = m_level[parent] + 1;  // root: 0xFFFFFFFF + 1 = 0
  m_level[idx] = lvl;
  if (m_levels.size() <= lvl) m_levels.resize(lvl + 1);
  m_levels[lvl].push_back(LevelEntry{parent, idx});
  return idx;
}

void TreeLayout::removeTransform(Index index)
{
  if (!isValidIndex(index)) throw std::runtime_error("Tree::removeTransform: Transform does not exist");
  auto& list = m_levels[m_level[index]];
  for (size_t e = 0; e < list.size(); ++e) {
    if (list[e].transform == index) {
      list[e] = list.back();
      list.pop_back();
      break;
    }
  }
  m_free[index >> 6] |= uint64_t(1) << (index & 63);
}

TreeLayout::Compiled TreeLayout::compile() const
{
  const size_t cap = m_level.size();
  Compiled c;
  c.index_of_slot.reserve(cap);
  c.slot_of_index.assign(cap, 0xFFFFFFFFu);
  c.parent_slot.assign(cap, 0);
  std::vector<uint32_t> parent_index;  // by slot
  parent_index.reserve(cap);
  c.index_of_slot.push_back(0);  // slot 0: the root
  parent_index.push_back(0);
  c.slot_of_index[0] = 0;
  c.level_offsets.push_back(1);

  // Per level: split into passes so that no pass both writes and reads one
  // index (sequential semantics of Tree.cpp:153-160 inside a level).
  std::vector<int> written(cap, -1), read(cap, -1);
  std::vector<uint32_t> touched;
  for (const auto& list : m_levels) {
    if (list.empty()) continue;
    std::vector<int> pass(list.size(), 0);
    int n_pass = 1;
    for (size_t e = 0; e < list.size(); ++e) {
      const uint32_t p = list[e].parent, t = list[e].transform;
      int s = 0;
      if (written[p] >= 0) s = std::max(s, written[p] + 1);  // reads a matrix written earlier in this level
      if (read[t] >= 0) s = std::max(s, read[t] + 1);        // writes a matrix read earlier in this level
      pass[e] = s;
      written[t] = s;
      read[p] = std::max(read[p], s);
      touched.push_back(p);
      touched.push_back(t);
      n_pass = std::max(n_pass, s + 1);
    }
    std::vector<size_t> order;
    for (int ps = 0; ps < n_pass; ++ps) {
      // Inside a pass no entry reads what another writes, so its order is free:
      // ordering it by parent slot makes the children of a slot range
      // contiguous (what tile runs of the propagation plan need).  Parents with
      // no slot yet (unlisted indices, placed after every level) sort last.
      order.clear();
      for (size_t e = 0; e < list.size(); ++e)
        if (pass[e] == ps) order.push_back(e);
      auto key = [&](size_t e) { return c.slot_of_index[list[e].parent]; };
      std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key(x) < key(y); });
      for (size_t e : order) {
        c.slot_of_index[list[e].transform] = static_cast<uint32_t>(c.index_of_slot.size());
        c.index_of_slot.push_back(list[e].transform);
        parent_index.push_back(list[e].parent);
      }
      c.level_offsets.push_back(c.index_of_slot.size());
    }
    for (uint32_t i : touched) written[i] = read[i] = -1;
    touched.clear();
  }
  // Every index in no list (free, or removed while children still read it) keeps a slot.
  for (size_t i = 0; i < cap; ++i) {
    if (c.slot_of_index[i] == 0xFFFFFFFFu) {
      c.slot_of_index[i] = static_cast<uint32_t>(c.index_of_slot.size());
      c.index_of_slot.push_back(static_cast<uint32_t>(i));
      parent_index.push_back(0);
    }
  }
  for (size_t s = 0; s < cap; ++s) c.parent_slot[s] = c.slot_of_index[parent_index[s]];
  c.parent_slot[0] = 0;
  return c;
}

// ---------------------------------------------------------------------------
// Tree
// ---------------------------------------------------------------------------
struct Tree::Device {
  void* stream = nullptr;
  size_t ld = 0;
  DevBuf<float> local, world, aos, it;
  DevBuf<uint32_t> parent_slot, index_of_slot, slot_of_index, old_slot, idx, bits;
  DevBuf<uint8_t> dirty_local, dirty_world;
  DevBuf<float> delta;
  dpc_tree_plan* plan = nullptr;  // launch schedule of the current layout
  ~Device()
  {
    if (plan) dpc_tree_plan_destroy(plan);
    if (stream) dpc_stream_destroy(stream);
  }
};

Tree::Tree() : m_dev(new Device)
{
  check(dpc_stream_create(&m_dev->stream), "dpc_stream_create");
  m_local.assign(m_layout.capacity(), Mat44f{});
  m_local[0] = Mat44f::identity();
  m_dirtyLocal.assign(m_layout.capacity(), 0);
  m_dirtyLocal[0] = 1;  // Tree.cpp:43
  m_worldHost.assign(m_layout.capacity(), Mat44f{});
  m_worldHost[0] = Mat44f::identity();
}

Tree::~Tree() = default;

void Tree::ensureCapacity()
{
  const size_t cap = m_layout.capacity();
  if (m_local.size() < cap) {
    m_local.resize(cap, Mat44f{});
    m_dirtyLocal.resize(cap, 0);
  }
}

Index Tree::addTransform(Index parentIndex, Mat44f const& matrix)
{
  const Index idx = m_layout.addTransform(parentIndex);
  ensureCapacity();
  m_local[idx] = matrix;
  m_dirtyLocal[idx] = 1;
  m_anyDirty = true;
  m_structureDirty = true;
  return idx;
}

void Tree::removeTransform(Index index)
{
  m_layout.removeTransform(index);
  m_structureDirty = true;
}

void Tree::updateLocalMatrix(Index index, Mat44f const& matrix)
{
  if (index >= m_local.size()) throw std::out_of_range("Tree::updateLocalMatrix: index beyond the tree");
  m_local[index] = matrix;
  m_dirtyLocal[index] = 1;
  m_anyDirty = true;
}

void Tree::recompile()
{
  // The new layout is built in buffers of its own and the tree takes it over
  // only when every step succeeded: a failure (an exception) leaves the tree
  // with its previous layout, worlds and plan, as if recompile had not run.
  Device& d = *m_dev;
  TreeLayout::Compiled next = m_layout.compile();
  const size_t cap = m_layout.capacity();
  const size_t old_ld = d.ld;
  const bool had_world = old_ld != 0;
  DevBuf<float> local, world;
  DevBuf<uint32_t> parent_slot, index_of_slot, slot_of_index, bits;
  DevBuf<uint8_t> dirty_local, dirty_world;
  local.reserve(4 * 4 * cap);
  world.reserve(4 * 4 * cap);
  parent_slot.reserve(cap);
  index_of_slot.reserve(cap);
  slot_of_index.reserve(cap);
  dirty_local.reserve(cap);
  dirty_world.reserve(cap + 4);
  bits.reserve(2 * ((cap + 63) / 64));
  d.aos.reserve(16 * cap);  // staging only: its contents are never kept
  d.old_slot.reserve(cap);

  // World matrices keep their values across the relayout: new slot s takes the
  // old slot of the same index (indices beyond the old capacity start at zero).
  std::vector<Mat44f> local_by_slot(cap);
  for (size_t s = 0; s < cap; ++s) local_by_slot[s] = m_local[next.index_of_slot[s]];
  std::vector<uint32_t> old_slot;  // host buffers stay alive until the stream is synchronised below
  std::vector<Mat44f> w;
  try {
    check(dpc_memcpy_h2d(d.aos.p, local_by_slot.data(), cap * sizeof(Mat44f), d.stream), "upload locals");
    check(dpc_matrices_from_aos(d.aos.p, 64, cap, local.p, cap, d.stream), "locals to planes");
    if (had_world) {
      old_slot.resize(cap);
      for (size_t s = 0; s < cap; ++s) {
        const uint32_t i = next.index_of_slot[s];
        old_slot[s] = i < m_compiled.slot_of_index.size() ? m_compiled.slot_of_index[i] : 0xFFFFFFFFu;
      }
      check(dpc_memcpy_h2d(d.old_slot.p, old_slot.data(), cap * sizeof(uint32_t), d.stream), "upload slot map");
      check(dpc_matrices_to_aos(d.world.p, old_ld, d.old_slot.p, cap, d.aos.p, 64, d.stream), "gather worlds");
    } else {
      w.assign(cap, Mat44f{});
      w[0] = Mat44f::identity();
      check(dpc_memcpy_h2d(d.aos.p, w.data(), cap * sizeof(Mat44f), d.stream), "upload worlds");
    }
    check(dpc_matrices_from_aos(d.aos.p, 64, cap, world.p, cap, d.stream), "worlds to planes");
    check(dpc_memcpy_h2d(parent_slot.p, next.parent_slot.data(), cap * sizeof(uint32_t), d.stream), "parents");
    check(dpc_memcpy_h2d(index_of_slot.p, next.index_of_slot.data(), cap * sizeof(uint32_t), d.stream), "slots");
    check(dpc_memcpy_h2d(slot_of_index.p, next.slot_of_index.data(), cap * sizeof(uint32_t), d.stream), "slots");
    check(dpc_fill_u32(reinterpret_cast<uint32_t*>(dirty_world.p), 0u, (cap + 4) / 4, d.stream), "clear dirty");
  } catch (...) {
    (void)dpc_stream_sync(d.stream);  // the host buffers above must outlive the copies already queued
    throw;
  }
  check(dpc_stream_sync(d.stream), "recompile sync");
  dpc_tree_plan* plan = nullptr;
  check(dpc_tree_plan_create(next.parent_slot.data(), cap, next.level_offsets.data(), next.level_offsets.size() - 1, 0,
                             &plan),
        "dpc_tree_plan_create");
  // commit: nothing below can fail
  if (d.plan) dpc_tree_plan_destroy(d.plan);
  d.plan = plan;
  d.local.swap(local);
  d.world.swap(world);
  d.parent_slot.swap(parent_slot);
  d.index_of_slot.swap(index_of_slot);
  d.slot_of_index.swap(slot_of_index);
  d.dirty_local.swap(dirty_local);
  d.dirty_world.swap(dirty_world);
  d.bits.swap(bits);
  d.ld = cap;
  m_compiled = std::move(next);
  m_structureDirty = false;
  m_itAll = true;  // slots moved: recompute every inverse transpose from the relaid-out worlds
}

void Tree::compute(Mat44f const&)
{
  TraceRange trace("Update TransformTree");
  Device& d = *m_dev;
  const size_t cap = m_layout.capacity();
  std::vector<uint32_t> slots;  // host buffers stay alive until the stream is synchronised below
  std::vector<Mat44f> mats;
  std::vector<uint8_t> dirty_by_slot;
  SyncOnUnwind sync(d.stream);  // ... or, when a step fails, until the copies already queued are done
  // A failure after the propagation was queued leaves its dirty-world bytes
  // set; they are cleared before the next propagation, so that its event holds
  // only the matrices it changes.
  struct StaleOnUnwind {
    bool& stale;
    int exceptions = std::uncaught_exceptions();
    ~StaleOnUnwind()
    {
      if (std::uncaught_exceptions() > exceptions) stale = true;
    }
  } mark_stale{m_dirtyWorldStale};
  if (m_structureDirty || d.ld != cap) {
    recompile();  // (every local uploaded; new dirty-world bytes, cleared)
    m_dirtyWorldStale = false;
  } else {
    if (m_dirtyWorldStale) {
      check(dpc_fill_u32(reinterpret_cast<uint32_t*>(d.dirty_world.p), 0u, (cap + 4) / 4, d.stream), "clear dirty");
      m_dirtyWorldStale = false;
    }
    if (m_anyDirty) {
      // upload only the changed local matrices
      for (size_t i = 0; i < cap; ++i) {
        if (m_dirtyLocal[i]) {
          slots.push_back(m_compiled.slot_of_index[i]);
          mats.push_back(m_local[i]);
        }
      }
      if (!slots.empty()) {
        d.idx.reserve(slots.size());
        d.delta.reserve(16 * slots.size());
        check(dpc_memcpy_h2d(d.idx.p, slots.data(), slots.size() * 4, d.stream), "upload delta index");
        check(dpc_memcpy_h2d(d.delta.p, mats.data(), mats.size() * 64, d.stream), "upload delta");
        check(dpc_matrices_scatter(d.local.p, d.ld, d.idx.p, d.delta.p, slots.size(), d.stream), "scatter delta");
      }
    }
  }
  dirty_by_slot.assign(cap, 0);
  for (size_t s = 0; s < cap; ++s) dirty_by_slot[s] = m_dirtyLocal[m_compiled.index_of_slot[s]];
  check(dpc_memcpy_h2d(d.dirty_local.p, dirty_by_slot.data(), cap, d.stream), "upload dirty");
  check(dpc_tree_propagate_plan(d.plan, d.local.p, d.world.p, d.dirty_local.p, d.dirty_world.p, d.stream),
        "dpc_tree_propagate_plan");
  if (m_itEnabled) updateInverseTranspose();
  // EventWorldMatricesChanged: the dirty-world bytes as a bitset over indices.
  const size_t words32 = 2 * ((cap + 63) / 64);
  check(dpc_fill_u32(d.bits.p, 0u, words32, d.stream), "clear bits");
  check(dpc_bytes_to_bits_indexed(d.dirty_world.p, d.index_of_slot.p, cap, d.bits.p, cap, d.stream), "event bits");
  m_lastEvent.assign((cap + 63) / 64, 0);
  check(dpc_memcpy_d2h(m_lastEvent.data(), d.bits.p, m_lastEvent.size() * 8, d.stream), "download event");
  check(dpc_fill_u32(reinterpret_cast<uint32_t*>(d.dirty_world.p), 0u, (cap + 4) / 4, d.stream), "clear dirty");
  check(dpc_stream_sync(d.stream), "compute sync");
  ++m_computeCount;
  for (TreeObserver* o : m_observers) o->onWorldMatricesChanged(m_lastEvent, cap);
  std::fill(m_dirtyLocal.begin(), m_dirtyLocal.end(), 0);
  m_anyDirty = false;
  m_worldHostStale = true;
  m_itHostStale = true;
}

// Runs on the tree's stream after propagation, while the dirty-world bytes
// still mark the matrices this compute changed.
void Tree::updateInverseTranspose()
{
  Device& d = *m_dev;
  if (m_itAll) {
    d.it.reserve(16 * d.ld);
    check(dpc_fill_u32(reinterpret_cast<uint32_t*>(d.it.p), 0u, 16 * d.ld, d.stream), "clear inverse transposes");
  }
  check(dpc_inverse_transpose(d.world.p, d.ld, m_itAll ? nullptr : d.dirty_world.p, d.ld, d.it.p, d.ld, nullptr,
                              d.stream),
        "dpc_inverse_transpose");
  m_itAll = false;
}

void Tree::setInverseTransposeEnabled(bool enabled)
{
  if (enabled && !m_itEnabled) {
    m_itAll = true;
    if (m_dev->ld != 0 && !m_structureDirty) {
      // worlds exist already: compute every inverse transpose now
      updateInverseTranspose();
      check(dpc_stream_sync(m_dev->stream), "inverse transpose sync");
    }
  }
  m_itEnabled = enabled;
  m_itHostStale = true;
}

Mat44f const* Tree::getWorldMatricesInverseTranspose() const
{
  if (!m_itEnabled) throw std::logic_error("Tree::getWorldMatricesInverseTranspose: not enabled");
  if (m_itHostStale) {
    Device& d = *m_dev;
    const size_t cap = m_layout.capacity();
    m_itHost.assign(cap, Mat44f{});
    if (d.ld == cap && d.it.p) {
      check(dpc_matrices_to_aos(d.it.p, d.ld, d.slot_of_index.p, cap, d.aos.p, 64, d.stream), "inverse to aos");
      check(dpc_memcpy_d2h(m_itHost.data(), d.aos.p, cap * 64, d.stream), "download inverse transposes");
      check(dpc_stream_sync(d.stream), "download sync");
    }
    m_itHostStale = false;
  }
  return m_itHost.data();
}

const float* Tree::deviceWorldInverseTransposePlanes() const { return m_itEnabled ? m_dev->it.p : nullptr; }

Mat44f const* Tree::getWorldMatrices() const
{
  if (m_worldHostStale) {
    Device& d = *m_dev;
    const size_t cap = m_layout.capacity();
    m_worldHost.resize(cap);
    if (d.ld == cap) {
      check(dpc_matrices_to_aos(d.world.p, d.ld, d.slot_of_index.p, cap, d.aos.p, 64, d.stream), "worlds to aos");
      check(dpc_memcpy_d2h(m_worldHost.data(), d.aos.p, cap * 64, d.stream), "download worlds");
      check(dpc_stream_sync(d.stream), "download sync");
    }
    m_worldHostStale = false;
  }
  return m_worldHost.data();
}

Mat44f const& Tree::getWorldMatrix(Index index) const { return getWorldMatrices()[index]; }

void Tree::attach(TreeObserver* observer) { m_observers.push_back(observer); }

void Tree::detach(TreeObserver* observer)
{
  m_observers.erase(std::remove(m_observers.begin(), m_observers.end(), observer), m_observers.end());
}

const float* Tree::deviceWorldPlanes() const { return m_dev->world.p; }
size_t Tree::deviceLd() const { return m_dev->ld; }
const uint32_t* Tree::deviceSlotOfIndex() const { return m_dev->slot_of_index.p; }
void* Tree::stream() const { return m_dev->stream; }

}  // namespace transform
}  // namespace dpcull

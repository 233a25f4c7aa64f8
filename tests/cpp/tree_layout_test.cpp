// tree_layout_test.cpp -- the host bookkeeping of the GPU tree (TreeLayout:
// index allocation, level lists, removal, slot compile) against the reference's
// dp::transform::Tree on the recorded trace.  CPU only.
#include <algorithm>
#include <cstdio>
#include <stdexcept>

#include "dpcull/Tree.hpp"
#include "tree_trace.hpp"

using dpcull::transform::TreeLayout;

static int failures = 0;
#define EXPECT(cond, ...)                     \
  do {                                        \
    if (!(cond)) {                            \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);               \
      std::printf("\n");                      \
      ++failures;                             \
    }                                         \
  } while (0)

static void check_compile(const TreeLayout& layout)
{
  const auto c = layout.compile();
  const size_t cap = layout.capacity();
  EXPECT(c.index_of_slot.size() == cap && c.slot_of_index.size() == cap, "compiled sizes");
  std::vector<uint8_t> seen(cap, 0);
  for (size_t s = 0; s < cap; ++s) {
    const uint32_t i = c.index_of_slot[s];
    EXPECT(i < cap && !seen[i], "slot map is not a permutation at slot %zu", s);
    if (i < cap) seen[i] = 1;
    EXPECT(c.slot_of_index[i] == s, "slot_of_index inverse at %zu", s);
  }
  EXPECT(c.index_of_slot[0] == 0, "root is slot 0");
  // every level entry sits in a pass slot range, after every pass of earlier
  // levels, and its parent slot holds its parent index
  size_t level_end = c.level_offsets.front();
  for (const auto& list : layout.levels()) {
    if (list.empty()) continue;
    size_t lo = cap, hi = 0;
    for (const auto& e : list) {
      const uint32_t s = c.slot_of_index[e.transform];
      EXPECT(c.index_of_slot[c.parent_slot[s]] == e.parent, "parent slot of %u", e.transform);
      lo = std::min<size_t>(lo, s);
      hi = std::max<size_t>(hi, s + 1);
    }
    EXPECT(lo == level_end, "level starts right after the previous one (%zu vs %zu)", lo, level_end);
    EXPECT(hi - lo == list.size(), "level slots are contiguous");
    level_end = hi;
  }
  EXPECT(c.level_offsets.back() == level_end, "passes end where the levels end");
}

static size_t pass_count(const TreeLayout& layout) { return layout.compile().level_offsets.size() - 1; }

int main(int argc, char** argv)
{
  const std::string path = argc > 1 ? argv[1] : "tests/golden/ref_tree_trace.bin";
  const TreeTrace t = read_tree_trace(path);
  TreeLayout layout;
  size_t adds = 0, removes = 0, throws = 0;
  for (const TreeOp& op : t.ops) {
    if (op.kind == 0) {
      ++adds;
      try {
        const uint32_t idx = layout.addTransform(op.a);
        EXPECT(op.b == idx, "addTransform(parent %u) gave index %u, the reference %u", op.a, idx, op.b);
      } catch (std::runtime_error const&) {
        ++throws;
        EXPECT(op.b == 0xFFFFFFFFu, "addTransform(parent %u) threw, the reference did not", op.a);
      }
    } else if (op.kind == 2) {
      ++removes;
      try {
        layout.removeTransform(op.a);
        EXPECT(op.b == 0, "removeTransform(%u) succeeded, the reference threw", op.a);
      } catch (std::runtime_error const&) {
        ++throws;
        EXPECT(op.b != 0, "removeTransform(%u) threw, the reference did not", op.a);
      }
    } else if (op.kind == 3) {
      const TreeCompute& c = t.computes[op.a];
      const auto& lv = layout.levels();
      EXPECT(lv.size() == c.levels.size(), "compute %u: %zu levels, the reference %zu", op.a, lv.size(),
             c.levels.size());
      for (size_t l = 0; l < std::min(lv.size(), c.levels.size()); ++l) {
        EXPECT(lv[l].size() == c.levels[l].size(), "compute %u level %zu size", op.a, l);
        for (size_t e = 0; e < std::min(lv[l].size(), c.levels[l].size()); ++e)
          EXPECT(lv[l][e].parent == c.levels[l][e].first && lv[l][e].transform == c.levels[l][e].second,
                 "compute %u level %zu entry %zu: (%u,%u) vs reference (%u,%u)", op.a, l, e, lv[l][e].parent,
                 lv[l][e].transform, c.levels[l][e].first, c.levels[l][e].second);
      }
      check_compile(layout);
    }
  }
  // Index reuse can put the writer of a matrix on the level of an orphan that
  // still reads it: a(L0) -> b(L1) -> c(L2); q(L1) under a; remove b; r = child of
  // q takes b's index at L2, after c in the level-2 list.  The reference's
  // sequential loop lets c read b's old matrix before r overwrites it, so the
  // compile must put c and r in different passes, c first.
  TreeLayout reuse;
  const uint32_t a = reuse.addTransform(0);
  const uint32_t b = reuse.addTransform(a);
  const uint32_t c = reuse.addTransform(b);
  const uint32_t q = reuse.addTransform(a);
  EXPECT(pass_count(reuse) == 3, "three levels, three passes (%zu)", pass_count(reuse));
  reuse.removeTransform(b);
  const uint32_t r = reuse.addTransform(q);
  EXPECT(r == b, "lowest free index is reused (%u vs %u)", r, b);
  check_compile(reuse);
  EXPECT(pass_count(reuse) == 4, "level 2 split into two passes (%zu passes)", pass_count(reuse));
  {
    const auto comp = reuse.compile();
    EXPECT(comp.slot_of_index[c] < comp.slot_of_index[r], "the reader runs in the earlier pass");
  }
  std::printf("tree layout: %zu adds, %zu removes, %zu reference exceptions reproduced, %zu computes, %d failures\n",
              adds, removes, throws, t.computes.size(), failures);
  return failures ? 1 : 0;
}

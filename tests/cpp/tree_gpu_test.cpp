// tree_gpu_test.cpp -- dpcull::transform::Tree (world matrices on the GPU)
// against the reference's dp::transform::Tree on the recorded trace: after
// every compute, every world matrix the reference recomputed is bit-identical
// and the changed-matrix event carries exactly the reference's bits.  Then a
// 200k-node tree with per-frame local updates against the oracle.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>

#include "dpcull/Tree.hpp"
#include "oracle.h"
#include "tree_trace.hpp"

using namespace dpcull;
using namespace dpcull::transform;

static int failures = 0;
#define EXPECT(cond, ...)                                \
  do {                                                   \
    if (!(cond)) {                                       \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);   \
      std::printf(__VA_ARGS__);                          \
      std::printf("\n");                                 \
      ++failures;                                        \
    }                                                    \
  } while (0)

struct Recorder : TreeObserver {
  std::vector<uint64_t> bits;
  size_t count = 0;
  int calls = 0;
  void onWorldMatricesChanged(const std::vector<uint64_t>& b, size_t c) override
  {
    bits = b;
    count = c;
    ++calls;
  }
  bool bit(size_t i) const { return (bits[i >> 6] >> (i & 63)) & 1; }
};

static Mat44f to_mat(const float* m)
{
  Mat44f r;
  std::memcpy(r.m, m, 64);
  return r;
}

static void replay_reference_trace(const std::string& path)
{
  const TreeTrace t = read_tree_trace(path);
  Tree tree;
  Recorder rec;
  tree.attach(&rec);
  for (const TreeOp& op : t.ops) {
    if (op.kind == 0) {
      try {
        const Index i = tree.addTransform(op.a, to_mat(op.m));
        EXPECT(i == op.b, "add gave %u, reference %u", i, op.b);
      } catch (std::runtime_error const&) {
        EXPECT(op.b == 0xFFFFFFFFu, "add(%u) threw", op.a);
      }
    } else if (op.kind == 1) {
      tree.updateLocalMatrix(op.a, to_mat(op.m));
    } else if (op.kind == 2) {
      try {
        tree.removeTransform(op.a);
      } catch (std::runtime_error const&) {
        EXPECT(op.b != 0, "remove(%u) threw", op.a);
      }
    } else if (op.kind == 3) {
      tree.compute();
      const TreeCompute& c = t.computes[op.a];
      const size_t nw = c.event.size();
      const Mat44f* w = tree.getWorldMatrices();
      size_t checked = 0, kept = 0;
      for (size_t i = 0; i < nw; ++i) {
        EXPECT(rec.bit(i) == (c.event[i] != 0), "compute %u: event bit %zu", op.a, i);
        // Every index below the reference's highest one has been allocated at
        // some point (lowest-free allocation), so its world matrix is defined:
        // recomputed now, or kept from an earlier compute across relayouts
        // (live transforms that did not change, and removed ones whose
        // orphaned children still read them).
        EXPECT(std::memcmp(w[i].m, &c.world[16 * i], 64) == 0, "compute %u: world matrix %zu differs (%s)", op.a, i,
               c.event[i] ? "recomputed" : "kept");
        ++(c.event[i] ? checked : kept);
      }
      for (size_t i = nw; i < rec.count; ++i) EXPECT(!rec.bit(i), "event bit beyond the reference's indices");
      std::printf("  reference compute %u: %zu recomputed and %zu kept matrices bit-identical\n", op.a, checked, kept);
    }
  }
  EXPECT(rec.calls == static_cast<int>(t.computes.size()), "one event per compute");
}

// A large random tree built depth first (as a scene traversal adds transforms),
// then frames of local updates; checked against orc_tree_compute on the same
// level lists.
static void random_frames()
{
  std::mt19937 rng(1234);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  auto rand_mat = [&]() {
    Mat44f m = Mat44f::identity();
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r][c] = (r == c ? 1.f : 0.f) + 0.3f * u(rng);
    for (int c = 0; c < 3; ++c) m[3][c] = 10.f * u(rng);
    return m;
  };
  Tree tree;
  std::vector<Mat44f> local(65536 * 4, Mat44f{});
  local[0] = Mat44f::identity();
  std::vector<Index> stack{0};
  std::vector<Index> nodes;
  const size_t target = 200000;
  while (nodes.size() < target) {
    // depth-first: add a child to the top of the stack, sometimes pop
    const Index parent = stack.back();
    const Mat44f m = rand_mat();
    const Index i = tree.addTransform(parent, m);
    local[i] = m;
    nodes.push_back(i);
    if (stack.size() < 14 && (rng() % 3 != 0)) stack.push_back(i);
    else if (stack.size() > 1 && rng() % 2) stack.pop_back();
  }
  const size_t cap = tree.getTransformCount();
  std::vector<float> world(16 * cap, 0.f), loc(16 * cap, 0.f);
  std::memcpy(world.data(), Mat44f::identity().m, 64);
  std::vector<uint8_t> dl(cap, 0), dw(cap, 0), ev(cap, 0);
  for (size_t i = 0; i < cap; ++i) std::memcpy(&loc[16 * i], local[i].m, 64);
  for (Index i : nodes) dl[i] = 1;
  const auto& levels = tree.layout().levels();
  std::vector<uint64_t> off{0};
  std::vector<uint32_t> entries;
  for (const auto& l : levels) {
    for (const auto& e : l) {
      entries.push_back(e.parent);
      entries.push_back(e.transform);
    }
    off.push_back(entries.size() / 2);
  }
  Recorder rec;
  tree.attach(&rec);
  tree.setInverseTransposeEnabled(true);
  for (int frame = 0; frame < 4; ++frame) {
    if (frame == 2) {  // off and on again with worlds present: every inverse transpose is recomputed at once
      tree.setInverseTransposeEnabled(false);
      tree.setInverseTransposeEnabled(true);
    }
    if (frame) {
      for (int k = 0; k < 500; ++k) {
        const Index i = nodes[rng() % nodes.size()];
        const Mat44f m = rand_mat();
        tree.updateLocalMatrix(i, m);
        std::memcpy(&loc[16 * i], m.m, 64);
        dl[i] = 1;
      }
    }
    tree.compute();
    orc_tree_compute(off.size() - 1, off.data(), entries.data(), loc.data(), world.data(), cap, dl.data(), dw.data(),
                     ev.data());
    const Mat44f* w = tree.getWorldMatrices();
    size_t bad = 0, changed = 0;
    for (Index i : nodes) {
      if (std::memcmp(w[i].m, &world[16 * i], 64) != 0) ++bad;
      if (ev[i] != rec.bit(i)) ++bad;
      changed += ev[i];
    }
    EXPECT(bad == 0, "frame %d: %zu world matrices or event bits differ", frame, bad);
    // inverse transposes of the world matrices (dp::math::invertTranspose), bit for bit
    const Mat44f* it = tree.getWorldMatricesInverseTranspose();
    size_t bad_it = 0;
    for (Index i : nodes) {
      float want[16];
      if (orc_invert_transpose(&world[16 * i], want) && std::memcmp(it[i].m, want, 64) != 0) ++bad_it;
    }
    EXPECT(bad_it == 0, "frame %d: %zu inverse transposes differ", frame, bad_it);
    std::printf("  frame %d: %zu nodes, %zu recomputed, %zu mismatches, %zu inverse-transpose mismatches\n", frame,
                nodes.size(), changed, bad, bad_it);
  }
}

// A tree holding only its root computes nothing and notifies an empty change
// set (Tree.cpp:133-166 visits no level); it must stay usable afterwards: add
// transforms, compute, remove them all (root only again), compute, add again.
static void root_only_tree()
{
  Tree tree;
  Recorder rec;
  tree.attach(&rec);
  auto none_set = [&]() {
    for (uint64_t w : rec.bits)
      if (w) return false;
    return true;
  };
  tree.compute();
  EXPECT(rec.calls == 1 && none_set(), "root-only compute: one empty event");
  EXPECT(std::memcmp(tree.getWorldMatrix(0).m, Mat44f::identity().m, 64) == 0, "root world is the identity");
  Mat44f a = Mat44f::identity(), b = Mat44f::identity();
  a[3][0] = 2.f;
  a[0][1] = 0.5f;
  b[3][2] = -3.f;
  b[1][1] = 1.25f;
  for (int round = 0; round < 2; ++round) {
    const Index ia = tree.addTransform(0, a);
    const Index ib = tree.addTransform(ia, b);
    tree.compute();
    // the oracle on the same two levels
    const size_t cap = tree.getTransformCount();
    std::vector<float> loc(16 * cap, 0.f), world(16 * cap, 0.f);
    std::vector<uint8_t> dl(cap, 0), dw(cap, 0), ev(cap, 0);
    std::memcpy(world.data(), Mat44f::identity().m, 64);
    std::memcpy(&loc[16 * ia], a.m, 64);
    std::memcpy(&loc[16 * ib], b.m, 64);
    dl[ia] = dl[ib] = 1;
    const uint64_t off[3] = {0, 1, 2};
    const uint32_t entries[4] = {0, ia, ia, ib};
    orc_tree_compute(2, off, entries, loc.data(), world.data(), cap, dl.data(), dw.data(), ev.data());
    const Mat44f* w = tree.getWorldMatrices();
    EXPECT(std::memcmp(w[ia].m, &world[16 * ia], 64) == 0 && std::memcmp(w[ib].m, &world[16 * ib], 64) == 0,
           "round %d: world matrices after adding to a root-only tree", round);
    EXPECT(rec.bit(ia) && rec.bit(ib) && !rec.bit(0), "round %d: event bits", round);
    tree.removeTransform(ib);
    tree.removeTransform(ia);
    tree.compute();
    EXPECT(none_set(), "round %d: root-only again: empty event", round);
  }
  EXPECT(rec.calls == 5, "one event per compute (%d)", rec.calls);
  std::printf("  root-only tree: %d computes\n", rec.calls);
}

int main(int argc, char** argv)
{
  const std::string path = argc > 1 ? argv[1] : "tests/golden/ref_tree_trace.bin";
  replay_reference_trace(path);
  root_only_tree();
  random_frames();
  std::printf("tree gpu: %d failures\n", failures);
  return failures ? 1 : 0;
}

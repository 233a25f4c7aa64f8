// tree_plan_test.cpp -- the propagation plan's tile runs (tree_plan.hpp),
// checked on the CPU.  Every tile run of every schedule must pass check_run,
// and an emulation of the plan -- tree_tile_kernel slot by slot from the
// run's pieces (memory reads see the arrays as they were before the launch,
// as in a launch that never reads what it writes), the other step kinds as
// sequential passes -- must give the world matrices and changed flags of a
// plain pass-by-pass propagation (dpc_tree_propagate's semantics, the order
// of dp::transform::Tree::compute, dp/transform/src/Tree.cpp:133-166) bit for
// bit, for full computes and for computes with a few dirty local matrices.
// Layouts come from TreeLayout (random trees, trees after removals and index
// reuse, adversarial shapes, and a 1,000,000-node tree shaped like the
// benchmark's).
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dpcull/Tree.hpp"
#include "tree_plan.hpp"

using dpcull::transform::TreeLayout;
namespace tp = dpc::tileplan;

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

struct M {
  float r[4][4];
};

// row r of L times the parent's rows, in the kernels' order (rowmat)
static M mul(const M& l, const M& p)
{
  M w;
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 4; ++c) {
      float x = l.r[i][0] * p.r[0][c];
      x = x + l.r[i][1] * p.r[1][c];
      x = x + l.r[i][2] * p.r[2][c];
      x = x + l.r[i][3] * p.r[3][c];
      w.r[i][c] = x;
    }
  return w;
}

struct Layout {
  std::vector<uint32_t> par;
  std::vector<uint64_t> off;
  size_t ld = 0;
};

static Layout compile(const TreeLayout& t)
{
  const auto c = t.compile();
  Layout L;
  L.par = c.parent_slot;
  L.off = c.level_offsets;
  L.ld = t.capacity();
  return L;
}

// pass-by-pass propagation of passes [a, b) (dpc_tree_propagate)
static void sequential(const Layout& lay, const std::vector<M>& loc, const std::vector<uint8_t>& dl, std::vector<M>& W,
                       std::vector<uint8_t>& DW, size_t a, size_t b)
{
  for (size_t k = a; k < b; ++k)
    for (uint64_t s = lay.off[k]; s < lay.off[k + 1]; ++s) {
      const uint64_t p = lay.par[s];
      if (p >= lay.ld) continue;
      if (!(dl[s] || DW[p])) continue;
      W[s] = mul(loc[s], W[p]);
      DW[s] = 1;
    }
}

// tree_tile_kernel over one run, tile by tile
static void emulate_run(const Layout& lay, const tp::Run& run, uint32_t threads, const std::vector<M>& loc,
                        const std::vector<uint8_t>& dl, std::vector<M>& W, std::vector<uint8_t>& DW)
{
  const std::vector<M> W_in = W;
  const std::vector<uint8_t> DW_in = DW;
  std::vector<uint8_t> written(lay.ld, 0);
  const uint32_t R = run.ring;
  for (size_t t = 0; t < run.tiles(); ++t) {
    const uint32_t g0 = run.tile_seg[t], n = run.tile_seg[t + 1] - g0;
    const tp::Seg* sg = run.segs.data() + g0;
    std::vector<M> rows(R);
    std::vector<uint8_t> flag(R, 0xEE);  // garbage until written
    uint32_t a = 0;
    while (a < n) {
      uint32_t b = a;
      while (true) {
        const bool end = (sg[b].flags & tp::kStepEnd) != 0;
        ++b;
        if (end || b >= n) break;
      }
      // each thread's piece and slot (load_step)
      // (a kQuad piece gives each slot four threads, one matrix row each: the
      // emulation computes the slot once, at its first thread)
      std::vector<uint32_t> mine(threads, UINT32_MAX), slot(threads, 0), row(threads, 0);
      for (uint32_t j = a; j < b; ++j) {
        const uint32_t width = (sg[j].flags & tp::kQuad) ? 4 : 1;
        for (uint32_t x = 0; x < threads; ++x)
          if (x - sg[j].pos < sg[j].count * width) {
            EXPECT(mine[x] == UINT32_MAX, "thread %u has two pieces in a step", x);
            mine[x] = j;
            slot[x] = sg[j].first + (x - sg[j].pos) / width;
            row[x] = (x - sg[j].pos) % width;
          }
      }
      for (uint32_t j = a; j < b; ++j) {
        const tp::Seg& s = sg[j];
        for (uint32_t x = 0; x < threads; ++x) {
          if (mine[x] != j || row[x] != 0) continue;
          const uint64_t sl = slot[x];
          const uint64_t p = lay.par[sl];
          const bool valid = p < lay.ld;
          const uint32_t d = static_cast<uint32_t>(p) - s.pfirst;
          const bool staged = valid && !(s.flags & tp::kFromMemory) && d < s.pcount;
          uint32_t i = s.pring + d;
          i = i >= R ? i - R : i;
          if (staged) EXPECT(i < R, "parent ring position %u >= %u", i, R);
          i = staged && i < R ? i : 0u;
          const uint32_t dp_lds = staged ? flag[i] : 0u;
          if (staged) EXPECT(dp_lds == 0 || dp_lds == 1, "parent flag read before it was written (slot %" PRIu64 ")", sl);
          bool dirty = false;
          M par{};
          if (dp_lds == 1) {
            par = rows[i];
            dirty = true;
          } else if (valid && (!staged || dl[sl])) {
            par = W_in[p];
            dirty = (dl[sl] | (staged ? 0 : DW_in[p])) != 0;
          }
          const bool kept = (s.flags & tp::kStaged) != 0;
          uint32_t o = 0;
          if (kept) {
            o = s.ring + static_cast<uint32_t>(sl - s.first);
            o = o >= R ? o - R : o;
            EXPECT(o < R, "ring position %u >= %u", o, R);
            if (o >= R) continue;
            flag[o] = dirty ? 1 : 0;
          }
          if (dirty) {
            const M w = mul(loc[sl], par);
            if (kept) rows[o] = w;
            if (sl < s.own_end) {
              EXPECT(!written[sl], "slot %" PRIu64 " written twice in one launch", sl);
              written[sl] = 1;
              W[sl] = w;
              DW[sl] = 1;
            }
          }
        }
      }
      a = b;
    }
  }
}

static std::string describe(const std::vector<tp::Step>& sched)
{
  std::string s;
  char buf[96];
  for (const auto& st : sched) {
    if (st.kind == tp::Step::kTiles)
      std::snprintf(buf, sizeof buf, "tiles(%zu-%zu: %zu, ring %u) ", st.l, st.r - 1, st.run.tiles(), st.run.ring);
    else if (st.kind == tp::Step::kSmall)
      std::snprintf(buf, sizeof buf, "small(%zu-%zu) ", st.l, st.r - 1);
    else
      std::snprintf(buf, sizeof buf, "level(%zu) ", st.l);
    s += buf;
  }
  return s;
}

struct Config {
  uint32_t threads, max_ring, ring_step;
  uint64_t capacity;
};

// Schedules the layout, checks every run, and compares the emulated plan with
// sequential propagation over `frames` computes (the first one full).
static void check_layout(const char* name, const Layout& lay, const Config& cfg, int frames, uint64_t seed,
                         bool verbose)
{
  const size_t n = lay.off.size() - 1;
  const auto sched = tp::schedule(lay.par.data(), lay.off.data(), n, cfg.threads, cfg.max_ring, cfg.capacity, 1024,
                                  32, 32, cfg.ring_step);
  size_t tiles = 0;
  for (const auto& st : sched) {
    if (st.kind != tp::Step::kTiles) continue;
    tiles += st.run.tiles();
    EXPECT(st.run.ring > 0 && st.run.ring <= cfg.max_ring, "%s: ring %u", name, st.run.ring);
    const std::string why = tp::check_run(lay.par.data(), lay.off.data(), tp::limits_for(cfg.threads, st.run.ring), st.run);
    EXPECT(why.empty(), "%s: run %zu-%zu: %s", name, st.l, st.r - 1, why.c_str());
  }
  // the schedule covers every pass once, in order
  size_t next = 0;
  for (const auto& st : sched) {
    while (next < n && lay.off[next + 1] == lay.off[next]) ++next;
    EXPECT(st.l == next, "%s: step starts at pass %zu, expected %zu", name, st.l, next);
    next = st.r;
  }
  while (next < n && lay.off[next + 1] == lay.off[next]) ++next;
  EXPECT(next == n, "%s: schedule ends at pass %zu of %zu", name, next, n);
  if (verbose)
    std::printf("%s: %zu slots, %zu passes, threads %u: %s(%zu tiles)\n", name, (size_t)lay.off[n], n, cfg.threads,
                describe(sched).c_str(), tiles);

  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> u(-2.f, 2.f);
  std::vector<M> loc(lay.ld);
  for (auto& m : loc)
    for (auto& r : m.r)
      for (auto& x : r) x = u(rng);
  std::vector<M> W_ref(lay.ld), W(lay.ld);
  for (size_t s = 0; s < lay.ld; ++s) {
    M id{};
    for (int i = 0; i < 4; ++i) id.r[i][i] = 1.f;
    W_ref[s] = W[s] = id;  // the root and unlisted slots
  }
  for (int frame = 0; frame < frames; ++frame) {
    std::vector<uint8_t> dl(lay.ld, frame == 0 ? 1 : 0);
    if (frame > 0) {
      const size_t k = 1 + rng() % std::max<size_t>(1, lay.ld / 50);
      for (size_t q = 0; q < k; ++q) {
        const size_t s = 1 + rng() % (lay.ld - 1);
        dl[s] = 1;
        for (auto& r : loc[s].r)
          for (auto& x : r) x = u(rng);
      }
    }
    std::vector<uint8_t> DW_ref(lay.ld, 0), DW(lay.ld, 0);
    sequential(lay, loc, dl, W_ref, DW_ref, 0, n);
    for (const auto& st : sched) {
      if (st.kind == tp::Step::kTiles)
        emulate_run(lay, st.run, cfg.threads, loc, dl, W, DW);
      else
        sequential(lay, loc, dl, W, DW, st.l, st.r);
    }
    size_t bad = 0, bad_flags = 0;
    for (size_t s = 1; s < lay.ld; ++s) {
      bad += std::memcmp(&W[s], &W_ref[s], sizeof(M)) != 0;
      bad_flags += DW[s] != DW_ref[s];
    }
    EXPECT(bad == 0 && bad_flags == 0, "%s frame %d: %zu matrices and %zu flags differ", name, frame, bad, bad_flags);
  }
}

// A random tree: `roots` children of the root, then breadth first each node
// gets min(min_children + failures before a success of probability p,
// max_children) children while depth < max_depth, until n nodes exist.
static void grow(TreeLayout& t, size_t n, uint64_t seed, double p, uint32_t max_children, uint32_t max_depth,
                 uint32_t roots, uint32_t min_children = 0)
{
  std::mt19937_64 rng(seed);
  std::geometric_distribution<uint32_t> geo(p);
  std::vector<std::pair<uint32_t, uint32_t>> q;  // (index, depth)
  size_t made = 0;
  for (uint32_t r = 0; r < roots && made < n; ++r, ++made) q.push_back({t.addTransform(0), 1});
  for (size_t h = 0; h < q.size() && made < n; ++h) {
    if (q[h].second >= max_depth) continue;
    const uint32_t c = std::min(geo(rng) + min_children, max_children);
    for (uint32_t k = 0; k < c && made < n; ++k, ++made) q.push_back({t.addTransform(q[h].first), q[h].second + 1});
  }
}

int main(int argc, char** argv)
{
  // --small: everything but the 1M-node tree (the sanitizer build runs this)
  const bool small = argc > 1 && std::string(argv[1]) == "--small";
  const Config real512{512, 1152, 64, 512}, real256{256, 576, 64, 1024}, real1024{1024, 2048, 64, 256};
  const Config tiny{32, 48, 8, 6}, tiny2{64, 96, 16, 3};

  // small random trees under small limits: many tiles, ring wrap, pieces split across steps
  for (uint64_t seed = 1; seed <= 60; ++seed) {
    TreeLayout t;
    grow(t, 50 + seed * 37, seed, 0.3 + 0.01 * (seed % 30), 6, 3 + seed % 10, 1 + seed % 5, seed % 3 == 0);
    const Layout lay = compile(t);
    char name[64];
    std::snprintf(name, sizeof name, "random seed %" PRIu64, seed);
    check_layout(name, lay, tiny, 3, seed, seed <= 3);
    check_layout(name, lay, tiny2, 2, seed + 100, false);
  }

  // removals and index reuse: orphans read freed (later re-used) indices, passes split
  for (uint64_t seed = 1; seed <= 30; ++seed) {
    TreeLayout t;
    grow(t, 400 + seed * 20, seed, 0.35, 5, 8, 3, 1);
    std::mt19937_64 rng(seed * 7);
    for (int k = 0; k < 60; ++k) {
      const uint32_t i = 1 + static_cast<uint32_t>(rng() % (t.capacity() - 1));
      if (t.isValidIndex(i) && i != 0) t.removeTransform(i);
    }
    for (int k = 0; k < 80; ++k) {
      uint32_t par = static_cast<uint32_t>(rng() % t.capacity());
      if (!t.isValidIndex(par)) par = 0;
      t.addTransform(par);
    }
    const Layout lay = compile(t);
    char name[64];
    std::snprintf(name, sizeof name, "reuse seed %" PRIu64, seed);
    check_layout(name, lay, tiny, 3, seed, seed == 17);
    check_layout(name, lay, real512, 2, seed, seed == 17);
  }

  // adversarial shapes
  {
    TreeLayout t;  // a broom: one deep chain, thousands of leaves under the top of it
    uint32_t top = t.addTransform(0), cur = top;
    for (int d = 0; d < 20; ++d) cur = t.addTransform(cur);
    for (int k = 0; k < 5000; ++k) t.addTransform(top);
    check_layout("broom", compile(t), tiny, 3, 5, true);
    check_layout("broom", compile(t), real512, 2, 5, true);
  }
  {
    TreeLayout t;  // a wide fan: 20,000 children of one node, each with one child
    const uint32_t hub = t.addTransform(0);
    std::vector<uint32_t> kids;
    for (int k = 0; k < 20000; ++k) kids.push_back(t.addTransform(hub));
    for (uint32_t k : kids) t.addTransform(k);
    check_layout("fan", compile(t), real256, 2, 6, true);
    check_layout("fan", compile(t), tiny2, 2, 6, false);
  }
  {
    TreeLayout t;  // many roots, all leaves
    for (int k = 0; k < 3000; ++k) t.addTransform(0);
    check_layout("flat", compile(t), real512, 2, 7, true);
  }

  // the benchmark's shape: 1,000,000 nodes, geometric children (p = 0.4, at most 8), depth <= 16
  if (!small) {
    TreeLayout t;
    grow(t, 1000000, 2, 0.4, 8, 16, 8, 1);
    const Layout lay = compile(t);
    check_layout("1M", lay, real512, 3, 2, true);
    check_layout("1M", lay, real256, 2, 3, true);
    check_layout("1M", lay, real1024, 2, 4, true);
  }

  std::printf("tree plan: %d failures\n", failures);
  return failures ? 1 : 0;
}

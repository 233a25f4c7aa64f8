// tree_trace.hpp -- reader of tests/golden/ref_tree_trace.bin, a trace of
// dp::transform::Tree operations with the reference's results at every compute
// (written by tests/golden/gen_ref_fixtures.py from the compiled reference).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct TreeOp {
  uint32_t kind;  // 0 add(parent=a) -> index b (0xFFFFFFFF: threw), 1 update(a), 2 remove(a) (b != 0: threw), 3 compute #a
  uint32_t a, b;
  float m[16];
};

struct TreeCompute {
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> levels;  // (parent, transform)
  std::vector<float> world;                                        // n_world x 16
  std::vector<uint8_t> event;                                      // n_world
};

struct TreeTrace {
  std::vector<TreeOp> ops;
  std::vector<TreeCompute> computes;
};

inline TreeTrace read_tree_trace(const std::string& path)
{
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  auto rd = [&](void* p, size_t n) {
    if (std::fread(p, 1, n, f) != n) throw std::runtime_error("truncated trace " + path);
  };
  auto u32 = [&]() {
    uint32_t v;
    rd(&v, 4);
    return v;
  };
  char magic[4];
  rd(magic, 4);
  if (std::memcmp(magic, "DPTT", 4) != 0 || u32() != 1) throw std::runtime_error("bad trace header");
  TreeTrace t;
  t.ops.resize(u32());
  for (auto& op : t.ops) {
    op.kind = u32();
    op.a = u32();
    op.b = u32();
    rd(op.m, 64);
  }
  t.computes.resize(u32());
  for (auto& c : t.computes) {
    std::vector<uint32_t> sizes(u32());
    for (auto& s : sizes) s = u32();
    for (uint32_t s : sizes) {
      std::vector<std::pair<uint32_t, uint32_t>> lvl(s);
      for (auto& e : lvl) {
        e.first = u32();
        e.second = u32();
      }
      c.levels.push_back(lvl);
    }
    const uint32_t nw = u32();
    c.world.resize(16 * size_t(nw));
    rd(c.world.data(), c.world.size() * 4);
    c.event.resize(nw);
    rd(c.event.data(), nw);
  }
  std::fclose(f);
  return t;
}

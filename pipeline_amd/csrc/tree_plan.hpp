// tree_plan.hpp -- the launch schedule of transform-tree propagation (host code).
//
// The layout (dpcull::transform::TreeLayout::compile) puts the slots of each
// pass contiguously, passes in order, and orders each pass by parent slot.
// A run of consecutive passes in which every parent of a pass lies in the
// previous pass ("linked" passes) is one launch of tree_tile_kernel:
//
//   - the run's forest is cut into tiles (workgroups) that are consecutive
//     intervals of its depth-first post-order; an interval is a contiguous
//     slot range of every pass of the run (build_run), so every slot is in
//     exactly one tile's owned range;
//   - a parent outside a tile's interval is the first slot of its pass after
//     the interval: the tile recomputes it as one extra slot (the same inputs
//     give the same result) and only the owning tile writes it, so tiles
//     share at most one slot per pass;
//   - a tile's slots, pass after pass, are cut into steps of `threads` slots,
//     one slot per thread; a step holds pieces ("segments") of one or more
//     passes, so small passes share a step and a step's loads are issued
//     together;
//   - a pass's results stay in LDS for the next pass of the run, in a ring of
//     `ring` positions per tile: two consecutive passes of a tile fit in it
//     together, so writing a pass never overwrites the rows the next pass
//     reads.  The last pass is not kept (nothing in the run reads it).
//
// Nothing the launch writes is read from memory within it: the first pass of
// a run reads parents written by earlier launches (the plan refuses a run
// whose first pass has a parent among the run's later passes), and a parent
// inside the run is read from LDS when it was recomputed, or from memory when
// it was not (then no tile writes it).
//
// Reference semantics: dp::transform::Tree::compute (dp/transform/src/Tree.cpp:133-166).
#pragma once

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace dpc {
namespace tileplan {

// One contiguous piece of one pass of one tile, handled within one step.
// 32 bytes: the kernel stages a tile's pieces in LDS and reads the first 16
// bytes (what finding a thread's piece needs) with one LDS load.
struct alignas(16) Seg {
  uint32_t first;    // first slot of the piece
  uint16_t count;    // slots in the piece
  uint16_t pos;      // first thread of the piece (a slot takes 1 thread, or 4 with kQuad)
  uint16_t ring;     // ring position of `first` (kept passes)
  uint16_t flags;
  uint32_t own_end;  // the tile writes this pass's slots below own_end to memory
  uint32_t pfirst;   // the tile's range of the previous pass: first slot,
  uint16_t pcount;   //   its length,
  uint16_t pring;    //   and the ring position of pfirst
  uint32_t tile_pieces;  // in a tile's first piece: the tile's pieces
  uint32_t tile_steps;   //   and steps (0 in the others)
};
static_assert(sizeof(Seg) == 32, "Seg is two 16-byte LDS loads");

enum : uint32_t {
  kStaged = 1,      // results go to the ring: the next pass of the run reads them
  kBarrier = 2,     // last piece of a kept pass whose rows other waves read: an LDS barrier follows it
  kStepEnd = 4,     // last piece of its step
  kFromMemory = 8,  // a piece of the run's first pass: parents come from memory
  kQuad = 16,       // four threads per slot, one matrix row each (passes of a tile with few slots)
};

// A pass of a tile with at most kQuadMax slots runs four threads per slot:
// the chain of small passes at the top of a tile then costs a quarter of the
// arithmetic latency per pass.
constexpr uint64_t kQuadMax = 16;

struct Limits {
  uint32_t threads;   // slots per step (threads per workgroup)
  uint32_t ring;      // LDS ring positions per tile
  uint32_t free_max;  // slots of the last pass per tile
  uint32_t max_segs;  // pieces per tile (the kernel's LDS table)
};

inline Limits limits_for(uint32_t threads, uint32_t ring)
{
  return Limits{threads, ring, threads * 4, 64};
}

struct Run {
  size_t l = 0, r = 0;             // passes [l, r)
  uint32_t ring = 0;               // LDS ring positions per tile
  std::vector<Seg> segs;           // every tile's pieces, tile after tile
  std::vector<uint32_t> tile_seg;  // tiles + 1 offsets into segs
  size_t tiles() const { return tile_seg.empty() ? 0 : tile_seg.size() - 1; }
};

// link[j] (j >= 1): pass j is nonempty, every parent of pass j lies in pass
// j - 1, and pass j is ordered by parent slot.
inline std::vector<uint8_t> links(const uint32_t* par, const uint64_t* off, size_t n)
{
  std::vector<uint8_t> link(n, 0);
  for (size_t j = 1; j < n; ++j) {
    bool ok = off[j + 1] > off[j] && off[j] > off[j - 1];
    for (uint64_t s = off[j]; ok && s < off[j + 1]; ++s) {
      ok = par[s] >= off[j - 1] && par[s] < off[j];
      if (ok && s > off[j]) ok = par[s] >= par[s - 1];
    }
    link[j] = ok;
  }
  return link;
}

// The first pass of a run may not read a slot the run writes (its later passes).
inline bool first_pass_reads_outside(const uint32_t* par, const uint64_t* off, size_t l, size_t r)
{
  for (uint64_t s = off[l]; s < off[l + 1]; ++s)
    if (par[s] >= off[l + 1] && par[s] < off[r]) return false;
  return true;
}

// Tiles of the linked passes [l, r), r - l >= 2.  False when a tile cannot
// meet the limits (then a shorter run may).
//
// Tiles are consecutive intervals of the run's forest in depth-first
// post-order (children before their parent, subtrees in slot order).  Because
// each pass is ordered by parent slot, every prefix of that order is a prefix
// of every pass, so an interval is a contiguous slot range of each pass (the
// slots the tile owns).  A parent outside the interval comes after it in
// post-order and is then the first slot of its pass after the interval (the
// next tile owns it): the tile computes it too, as one extra slot at the end
// of its range in that pass.  The greedy cut closes a tile before the node
// that would break a limit.
inline bool build_run(const uint32_t* par, const uint64_t* off, size_t l, size_t r, const Limits& lim, Run& out)
{
  out = Run{};
  out.l = l;
  out.r = r;
  out.ring = lim.ring;
  out.tile_seg.assign(1, 0);
  if (r < l + 2 || !first_pass_reads_outside(par, off, l, r)) return false;
  if (lim.threads == 0 || lim.threads > 0xFFFF || lim.ring == 0 || lim.ring > 0xFFFF) return false;
  const size_t P = r - l;  // passes of the run; pass P-1 (the last) is not kept in LDS
  // first child of every slot of passes l..r-2 (children of a slot are contiguous)
  std::vector<std::vector<uint64_t>> fc(P - 1);
  for (size_t k = 0; k + 1 < P; ++k) {
    const uint64_t a = off[l + k], b = off[l + k + 1];
    fc[k].resize(b - a + 1);
    uint64_t c = off[l + k + 1];
    for (uint64_t v = a; v <= b; ++v) {
      while (c < off[l + k + 2] && par[c] < v) ++c;
      fc[k][v - a] = c;
    }
  }
  // the tile [c, C) (per pass: slots [c[k], C[k]) owned) -> range lengths with
  // the extra parent slots; false when a limit breaks
  std::vector<uint64_t> c(P), C(P), cnt(P);
  auto measure = [&](const std::vector<uint64_t>& hi) {
    uint64_t need = UINT64_MAX;  // slot the pass below needs from this pass (its last slot's parent)
    size_t passes = 0;
    uint64_t total = 0;
    for (size_t k = P; k-- > 0;) {
      uint64_t end = hi[k] > c[k] ? hi[k] : c[k];  // exclusive end of the range
      if (need != UINT64_MAX && need + 1 > end) end = need + 1;
      cnt[k] = end - c[k];
      need = (cnt[k] && k > 0) ? par[end - 1] : UINT64_MAX;
      if (cnt[k]) {
        ++passes;
        total += cnt[k] * (cnt[k] <= kQuadMax ? 4 : 1) + 3;  // threads, and at most 3 left over at a step's end
      }
    }
    for (size_t k = 0; k + 1 < P; ++k) {
      if (cnt[k] > lim.ring) return false;
      if (k && cnt[k - 1] + cnt[k] > lim.ring) return false;
    }
    if (cnt[P - 1] > lim.free_max) return false;
    // pieces: at most one per nonempty pass plus one per step boundary (threads counted per step)
    return passes + (total + lim.threads - 1) / lim.threads <= lim.max_segs;
  };
  auto emit = [&]() {
    std::vector<uint64_t> o(P, 0);  // ring positions of the kept passes
    for (size_t k = 1; k + 1 < P; ++k) o[k] = (o[k - 1] + cnt[k - 1]) % lim.ring;
    // Lay the pieces out over the tile's threads, step after step: a slot of
    // pass k takes width(k) threads and never straddles two steps.
    struct Piece {
      size_t k;
      uint64_t q, n, lane;  // slots [q, q + n) of the tile's range in pass k, from flattened thread `lane`
      bool step_end;
    };
    std::vector<Piece> pcs;
    const uint64_t T = lim.threads;
    auto round_up = [&](uint64_t x) { return (x + T - 1) / T * T; };
    uint64_t p = 0;
    for (size_t k = 0; k < P; ++k) {
      if (!cnt[k]) continue;
      const uint64_t w = cnt[k] <= kQuadMax ? 4 : 1;
      if (p % T + w > T) {  // the step has no room for one slot: close it
        pcs.back().step_end = true;
        p = round_up(p);
      }
      for (uint64_t q = 0; q < cnt[k];) {
        const uint64_t n = std::min<uint64_t>(cnt[k] - q, (T - p % T) / w);
        pcs.push_back(Piece{k, q, n, p, false});
        q += n;
        p += n * w;
        if (p % T == 0 || q < cnt[k]) {  // the step is full
          pcs.back().step_end = true;
          p = round_up(p);
        }
      }
    }
    pcs.back().step_end = true;
    // the wave (64-thread block of one step) holding all of pass k, or -1
    std::vector<int64_t> wave(P, -1);
    std::vector<uint64_t> lo(P, UINT64_MAX), hi(P, 0);
    for (const Piece& pc : pcs) {
      const uint64_t w = cnt[pc.k] <= kQuadMax ? 4 : 1;
      lo[pc.k] = std::min(lo[pc.k], pc.lane);
      hi[pc.k] = std::max(hi[pc.k], pc.lane + pc.n * w - 1);
    }
    for (size_t k = 0; k < P; ++k)
      if (cnt[k] && lo[k] / 64 == hi[k] / 64) wave[k] = static_cast<int64_t>(lo[k] / 64);
    for (const Piece& pc : pcs) {
      const size_t k = pc.k;
      Seg sg{};
      sg.first = static_cast<uint32_t>(c[k] + pc.q);
      sg.count = static_cast<uint16_t>(pc.n);
      sg.pos = static_cast<uint16_t>(pc.lane % T);
      sg.ring = k + 1 < P ? static_cast<uint16_t>((o[k] + pc.q) % lim.ring) : uint16_t(0);
      sg.own_end = static_cast<uint32_t>(C[k]);
      uint32_t flags = cnt[k] <= kQuadMax ? kQuad : 0u;
      if (k == 0) {
        flags |= kFromMemory;
      } else {
        sg.pfirst = static_cast<uint32_t>(c[k - 1]);
        sg.pcount = static_cast<uint16_t>(cnt[k - 1]);
        sg.pring = static_cast<uint16_t>(o[k - 1]);
      }
      if (k + 1 < P) {
        flags |= kStaged;
        // a barrier closes the pass unless its rows are written and read (by
        // the next pass) within one wave, whose LDS operations execute in
        // order; a pass that nothing reads needs none
        const bool readers = cnt[k + 1] > 0;
        const bool one_wave = wave[k] >= 0 && wave[k] == wave[k + 1];
        if (pc.q + pc.n == cnt[k] && readers && !one_wave) flags |= kBarrier;
      }
      if (pc.step_end) flags |= kStepEnd;
      sg.flags = static_cast<uint16_t>(flags);
      out.segs.push_back(sg);
    }
    const size_t first = out.tile_seg.back();
    uint32_t steps = 0;
    for (size_t j = first; j < out.segs.size(); ++j) steps += (out.segs[j].flags & kStepEnd) ? 1u : 0u;
    out.segs[first].tile_pieces = static_cast<uint32_t>(out.segs.size() - first);
    out.segs[first].tile_steps = steps;
    out.tile_seg.push_back(static_cast<uint32_t>(out.segs.size()));
  };
  for (size_t k = 0; k < P; ++k) c[k] = C[k] = off[l + k];
  std::vector<uint64_t> trial(P);
  bool open = false;  // the current tile owns at least one slot
  // iterative post-order over the forest rooted at pass l
  struct Frame {
    uint64_t v;
    size_t k;
    uint64_t next;  // next child to visit
  };
  std::vector<Frame> stack;
  for (uint64_t root = off[l]; root < off[l + 1]; ++root) {
    stack.push_back(Frame{root, 0, root + 0});
    stack.back().next = P > 1 ? fc[0][root - off[l]] : 0;
    while (!stack.empty()) {
      Frame& f = stack.back();
      if (f.k + 1 < P && f.next < fc[f.k][f.v - off[l + f.k] + 1]) {
        const uint64_t child = f.next++;
        const size_t ck = f.k + 1;
        stack.push_back(Frame{child, ck, ck + 1 < P ? fc[ck][child - off[l + ck]] : 0});
        continue;
      }
      // visit f.v: extend the tile by it, or close the tile before it
      const size_t k = f.k;
      trial = C;
      trial[k] = f.v + 1;
      if (open && !measure(trial)) {
        measure(C);
        emit();
        c = C;
      }
      C = trial;
      if (!measure(C)) return false;  // one slot and its ancestors do not fit
      open = true;
      stack.pop_back();
    }
  }
  if (open) {
    measure(C);
    emit();
  }
  return true;
}

// Checks every property tree_tile_kernel relies on: pieces inside their
// pass, the tile's range and the step; ring positions in range; the parent
// of every slot found at the ring position the tile wrote it to, and not
// overwritten by the pass that reads it; a kept pass without a closing
// barrier written and read (by the next pass) within one wave of one step;
// every slot of the run written by exactly one tile; the first pass reading
// nothing the run writes.  Returns an empty string, or the first violation.
inline std::string check_run(const uint32_t* par, const uint64_t* off, const Limits& lim, const Run& run)
{
  char msg[256];
  auto bad = [&](const char* what, size_t t, size_t j) {
    std::snprintf(msg, sizeof msg, "tile %zu piece %zu: %s", t, j, what);
    return std::string(msg);
  };
  const size_t l = run.l, r = run.r;
  if (r < l + 2) return "run shorter than two passes";
  if (!first_pass_reads_outside(par, off, l, r)) return "first pass reads a slot the run writes";
  if (run.tile_seg.empty() || run.tile_seg[0] != 0 || run.tile_seg.back() != run.segs.size())
    return "bad tile offsets";
  const uint64_t base = off[l], end = off[r];
  std::vector<uint8_t> owned(end - base, 0);
  std::vector<int64_t> ring_slot(lim.ring), ring_pass(lim.ring);
  for (size_t t = 0; t + 1 < run.tile_seg.size(); ++t) {
    const uint32_t g0 = run.tile_seg[t], g1 = run.tile_seg[t + 1];
    if (g1 <= g0) return bad("empty tile", t, 0);
    if (g1 - g0 > lim.max_segs) return bad("too many pieces", t, 0);
    if (!(run.segs[g1 - 1].flags & kStepEnd)) return bad("tile does not end a step", t, g1 - 1);
    uint32_t steps = 0;
    for (uint32_t j = g0; j < g1; ++j) steps += (run.segs[j].flags & kStepEnd) ? 1u : 0u;
    if (run.segs[g0].tile_pieces != g1 - g0 || run.segs[g0].tile_steps != steps)
      return bad("tile piece or step count wrong", t, g0);
    std::fill(ring_slot.begin(), ring_slot.end(), -1);
    std::fill(ring_pass.begin(), ring_pass.end(), -1);
    // pass index and wave block (step * threads + position) / 64 of every piece
    std::vector<size_t> pass_of(g1 - g0);
    std::vector<int64_t> wave_lo(g1 - g0), wave_hi(g1 - g0);
    size_t pass = l;
    uint64_t step = 0, next_pos = 0;
    uint64_t pass_first = 0, pass_count = 0, pass_ring = 0, pass_own = 0;
    uint64_t prev_first = 0, prev_count = 0, prev_ring = 0;
    bool have_pass = false;
    std::vector<uint32_t> written;  // ring positions written by the current pass
    auto commit = [&]() {           // the current pass's rows become readable
      for (size_t i = 0; i < written.size(); ++i) {
        ring_slot[written[i]] = static_cast<int64_t>(pass_first + i);
        ring_pass[written[i]] = static_cast<int64_t>(pass);
      }
      written.clear();
    };
    for (uint32_t j = g0; j < g1; ++j) {
      const Seg& s = run.segs[j];
      const uint64_t lanes = static_cast<uint64_t>(s.count) * ((s.flags & kQuad) ? 4 : 1);
      if (s.count == 0) return bad("empty piece", t, j);
      if (s.pos != next_pos) return bad("pieces of a step not contiguous", t, j);
      if (s.pos + lanes > lim.threads) return bad("piece beyond the step", t, j);
      wave_lo[j - g0] = static_cast<int64_t>((step * lim.threads + s.pos) / 64);
      wave_hi[j - g0] = static_cast<int64_t>((step * lim.threads + s.pos + lanes - 1) / 64);
      if (s.flags & kStepEnd) {
        next_pos = 0;
        ++step;
      } else {
        next_pos = s.pos + lanes;
      }
      // which pass: continue the current one or start the next
      const bool cont = have_pass && s.first == pass_first + pass_count;
      const bool in_cur = cont && s.first >= off[pass] && s.first < off[pass + 1];
      if (!in_cur) {
        if (have_pass) {
          commit();
          prev_first = pass_first;
          prev_count = pass_count;
          prev_ring = pass_ring;
          ++pass;
        }
        if (pass >= r) return bad("more passes than the run", t, j);
        have_pass = true;
        pass_first = s.first;
        pass_count = 0;
        pass_ring = s.ring;
        pass_own = s.own_end;
      }
      pass_of[j - g0] = pass;
      if (s.first < off[pass] || static_cast<uint64_t>(s.first) + s.count > off[pass + 1])
        return bad("piece outside its pass", t, j);
      if (s.own_end != pass_own) return bad("ownership changes within a pass", t, j);
      const bool kept = pass + 1 < r;
      if (kept != ((s.flags & kStaged) != 0)) return bad("kept flag wrong", t, j);
      if (((s.flags & kFromMemory) != 0) != (pass == l)) return bad("memory flag wrong", t, j);
      if (kept && s.ring != (pass_ring + pass_count) % lim.ring) return bad("ring positions not contiguous", t, j);
      if (pass > l) {
        if (s.pfirst != prev_first || s.pcount != prev_count || s.pring != prev_ring)
          return bad("parent range differs from the previous pass", t, j);
        if (s.pcount > lim.ring || s.pring >= lim.ring) return bad("parent range beyond the ring", t, j);
      }
      if (kept && s.ring >= lim.ring) return bad("ring position out of range", t, j);
      for (uint32_t q = 0; q < s.count; ++q) {
        const uint64_t slot = static_cast<uint64_t>(s.first) + q;
        if (pass > l) {
          const uint64_t p = par[slot];
          if (p < s.pfirst || p >= static_cast<uint64_t>(s.pfirst) + s.pcount) return bad("parent outside the tile", t, j);
          const uint64_t at = (s.pring + (p - s.pfirst)) % lim.ring;
          if (ring_slot[at] != static_cast<int64_t>(p)) return bad("parent not at its ring position", t, j);
        }
        if (slot < s.own_end) {
          if (owned[slot - base]) return bad("slot written by two tiles", t, j);
          owned[slot - base] = 1;
        }
        if (kept) {
          // the position may not hold a row of the previous pass, which this pass reads
          const uint32_t at = static_cast<uint32_t>((s.ring + q) % lim.ring);
          if (pass > l && ring_pass[at] == static_cast<int64_t>(pass - 1))
            return bad("pass overwrites the ring positions its parents occupy", t, j);
          written.push_back(at);
        }
      }
      pass_count += s.count;
      if (kept && pass_count > lim.ring) return bad("pass longer than the ring", t, j);
    }
    // a kept pass whose last piece has no barrier: it and the next pass lie in one wave
    for (uint32_t j = g0; j < g1; ++j) {
      const size_t k = pass_of[j - g0];
      const bool last_of_pass = j + 1 == g1 || pass_of[j + 1 - g0] != k;
      if (!last_of_pass || k + 1 >= r || (run.segs[j].flags & kBarrier)) continue;
      int64_t lo = INT64_MAX, hi = INT64_MIN;
      bool read = false;  // the tile has pieces of pass k + 1
      for (uint32_t i = g0; i < g1; ++i)
        if (pass_of[i - g0] == k || pass_of[i - g0] == k + 1) {
          read = read || pass_of[i - g0] == k + 1;
          lo = std::min(lo, wave_lo[i - g0]);
          hi = std::max(hi, wave_hi[i - g0]);
        }
      if (read && lo != hi) return bad("pass without a barrier is read by another wave", t, j);
    }
  }
  for (uint64_t s = base; s < end; ++s)
    if (!owned[s - base]) {
      std::snprintf(msg, sizeof msg, "slot %llu written by no tile", static_cast<unsigned long long>(s));
      return msg;
    }
  return std::string();
}

// A launch of the schedule: a tile run, a run of small passes in one
// workgroup, or one pass.
struct Step {
  enum Kind { kTiles, kSmall, kLevel } kind;
  size_t l, r;  // passes [l, r)
  Run run;      // kTiles
};

// The schedule for passes [0, n): tile runs wherever linked passes allow one
// (longest first), else runs of small passes (at most small_max slots
// together) in one workgroup, else one launch per pass.  Empty passes are
// skipped.  A tile run's ring is the smallest multiple of ring_step positions
// (up to max_ring, itself a multiple) that keeps its tiles within `capacity`
// (the workgroups the device holds at once), so that every tile of the run is
// resident together.
inline std::vector<Step> schedule(const uint32_t* par, const uint64_t* off, size_t n, uint32_t threads,
                                  uint32_t max_ring, uint64_t capacity, uint64_t small_max, size_t small_levels_max,
                                  size_t run_levels_max, uint32_t ring_step = 64)
{
  std::vector<Step> steps;
  const std::vector<uint8_t> link = links(par, off, n);
  size_t l = 0;
  while (l < n) {
    if (off[l + 1] == off[l]) {
      ++l;
      continue;
    }
    size_t rmax = l + 1;
    while (rmax < n && link[rmax] && rmax - l < run_levels_max) ++rmax;
    bool done = false;
    for (size_t r = rmax; r >= l + 2 && !done; --r) {
      Step st{Step::kTiles, l, r, Run{}};
      if (!build_run(par, off, l, r, limits_for(threads, max_ring), st.run)) continue;
      // the smallest ring that still fits the device in one wave of tiles
      uint32_t lo = 1, hi = max_ring / ring_step;  // ring = ring_step * k, k in [lo, hi]; hi is feasible
      if (st.run.tiles() <= capacity) {
        while (lo < hi) {
          const uint32_t mid = (lo + hi) / 2;
          Run trial;
          if (build_run(par, off, l, r, limits_for(threads, ring_step * mid), trial) && trial.tiles() <= capacity)
            hi = mid;
          else
            lo = mid + 1;
        }
        if (ring_step * hi != max_ring) build_run(par, off, l, r, limits_for(threads, ring_step * hi), st.run);
      }
      steps.push_back(std::move(st));
      l = r;
      done = true;
    }
    if (done) continue;
    if (off[l + 1] - off[l] <= small_max) {
      Step st{Step::kSmall, l, l, Run{}};
      while (st.r < n && st.r - l < small_levels_max && off[st.r + 1] - off[l] <= small_max) ++st.r;
      l = st.r;
      steps.push_back(std::move(st));
      continue;
    }
    steps.push_back(Step{Step::kLevel, l, l + 1, Run{}});
    ++l;
  }
  return steps;
}

}  // namespace tileplan
}  // namespace dpc

// Tile-shape selection for the persistent grouped GEMM.
//
// A contraction step hands the GEMM a list of tasks whose shapes follow the
// quantum-number block sizes of the tensors involved.  The autotuner reduces
// that list to a block-size-histogram key (flop-weighted mean log2 of M, N
// and K in half-octave bins, the number of tasks and the total work in
// octaves, the precision) and caches the fastest tile shape per key.  On a
// miss it times every shape on the live operands with HIP events: the launch
// is repeated, so this is done only when repeating it is harmless (beta = 0
// everywhere and no output overlaps an input); otherwise the heuristic
// choose_gemm_cfg decides.  Every shape computes each output element with the
// same K order, so the choice changes timing only, never the result bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "dmrj.h"
#include "internal.h"

namespace dmrj {

#define HIP_OK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw Error(-3, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

const TileShape kGemmShapes[kNumGemmShapes] = {{32, 32}, {64, 64}, {128, 128}, {64, 128}, {128, 64}};

static int occupancy(Context& c, int cfg, int arith) {
  const bool planes = arith == 2 && c.f16_presplit;
  int& o = c.occ_cache[planes ? 3 : arith][cfg];
  if (o == 0)
    o = arith == 1 ? gemm_occupancy_lp(cfg)
                   : (arith == 2 ? (planes ? gemm_occupancy_f16p(cfg) : gemm_occupancy_f16(cfg)) : gemm_occupancy_f64(cfg));
  return o;
}

// FP16x3 operand planes of a launch: one region per distinct operand block
// (the same pointer, leading dimension, shape, storage order and alpha), its
// planes at 256-byte aligned offsets of one transient buffer, the split work
// items and every segment's plane pointers.
struct F16PlaneOffsets {
  size_t regionmaxwork = 0, regions = 0, splitwork = 0, f16segs = 0;
};
static F16PlaneOffsets prepare_f16_planes(Context& c, GemmLaunch& g, const std::vector<GemmTask>& tasks,
                                          const std::vector<GemmSeg>& segs, UploadPack& up) {
  struct Key {
    const double* p;
    int64_t ld;
    int rows, K, tr, op;
    double alpha;
    bool operator<(const Key& o) const {
      if (p != o.p) return p < o.p;
      if (ld != o.ld) return ld < o.ld;
      if (rows != o.rows) return rows < o.rows;
      if (K != o.K) return K < o.K;
      if (tr != o.tr) return tr < o.tr;
      if (op != o.op) return op < o.op;
      return alpha < o.alpha;
    }
  };
  std::map<Key, int> idx;
  std::vector<F16Region> regs;
  std::vector<int64_t> off;
  std::vector<int2> rsel(segs.size(), make_int2(-1, -1));
  int64_t halves = 0;
  auto region = [&](const Key& k) {
    auto it = idx.find(k);
    if (it != idx.end()) return it->second;
    const int kpad = (k.K + 31) / 32 * 32;
    F16Region r{k.p, k.ld, k.rows, k.K, kpad, k.tr, k.op, k.alpha, nullptr, nullptr};
    regs.push_back(r);
    off.push_back(halves);
    halves += ((int64_t)2 * k.rows * kpad + 127) / 128 * 128;  // hi then lo, 256-byte aligned
    idx[k] = (int)regs.size() - 1;
    return (int)regs.size() - 1;
  };
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    const GemmTask& t = tasks[ti];
    for (int q = 0; q < t.nseg; ++q) {
      const int si = t.seg0 + q;
      const GemmSeg& sg = segs[si];
      if (sg.K <= 0 || t.M <= 0 || t.N <= 0) continue;
      const bool tA = sg.flags & 1, tB = sg.flags & 2;
      // A element (m, k): tA ? A[k + m lda] : A[m + k lda]; B element (k, n) as row n:
      // tB ? B[n + k ldb] : B[k + n ldb]
      rsel[si].x = region(Key{sg.A, sg.lda, t.M, sg.K, tA ? 1 : 0, 0, sg.alpha});
      rsel[si].y = region(Key{sg.B, sg.ldb, t.N, sg.K, tB ? 0 : 1, 1, 1.0});
    }
  }
  g.presplit = true;
  uint16_t* planes = (uint16_t*)c.alloc.alloc(std::max<int64_t>(halves, 128) * sizeof(uint16_t));
  c.transient.push_back(planes);  // lives until the next synchronisation point
  std::vector<int4> work, mwork;
  for (size_t r = 0; r < regs.size(); ++r) {
    regs[r].hi = planes + off[r];
    regs[r].lo = regs[r].hi + (int64_t)regs[r].rows * regs[r].kpad;
    for (int r0 = 0; r0 < regs[r].rows; r0 += 64)
      for (int k0 = 0; k0 < regs[r].kpad; k0 += 32) work.push_back(make_int4((int)r, r0, k0, 0));
    // the maximum pass walks the operand as stored: columns of srows elements
    const int64_t srows = regs[r].tr ? regs[r].K : regs[r].rows;
    const int64_t scols = regs[r].tr ? regs[r].rows : regs[r].K;
    const int64_t step = std::max<int64_t>(1, 16384 / std::max<int64_t>(srows, 1));
    for (int64_t c0 = 0; c0 < scols; c0 += step)
      mwork.push_back(make_int4((int)r, (int)c0, (int)std::min(scols, c0 + step), 0));
  }
  std::vector<F16Seg> fs(std::max<size_t>(1, segs.size()));
  for (size_t si = 0; si < segs.size(); ++si) {
    if (rsel[si].x < 0) {
      fs[si] = F16Seg{nullptr, nullptr, nullptr, nullptr, 0, 0, 0, 0};
      continue;
    }
    const F16Region& a = regs[rsel[si].x];
    const F16Region& b = regs[rsel[si].y];
    fs[si] = F16Seg{a.hi, a.lo, b.hi, b.lo, a.kpad, rsel[si].x, rsel[si].y, 0};
  }
  g.nregions = (int)regs.size();
  g.regionmax = (double*)c.alloc.alloc(std::max<size_t>(1, regs.size()) * sizeof(double));
  c.transient.push_back(g.regionmax);
  F16PlaneOffsets o;
  o.regionmaxwork = up.add(mwork.data(), mwork.size() * sizeof(int4));
  g.nregionmaxwork = (int)mwork.size();
  o.regions = up.add(regs.data(), regs.size() * sizeof(F16Region));
  o.splitwork = up.add(work.data(), work.size() * sizeof(int4));
  g.nsplitwork = (int)work.size();
  o.f16segs = up.add(fs.data(), fs.size() * sizeof(F16Seg));
  return o;
}

// XCD-aware order.  Work-group b of a launch runs on XCD b mod 8 (the
// dispatcher deals work-groups round-robin over the eight XCDs, each with its
// own L2), and the persistent grid (a multiple of 8, or one work-group per
// tile) gives tile t to work-group t mod grid, i.e. to XCD t mod 8.  Tiles
// that share a B panel (same task, same column block) form a group; groups
// are dealt to eight streams by least accumulated work in descending-work
// order, and position 8i + x of the launch order takes the i-th tile of
// stream x, so each XCD's L2 sees whole panel groups.
static std::vector<GemmTile> xcd_interleave(const std::vector<GemmTile>& tiles, const std::vector<GemmTask>& tasks,
                                            const std::vector<GemmSeg>& segs) {
  constexpr int kXcd = 8;
  std::vector<std::vector<GemmTile>> stream(kXcd);
  std::vector<double> load(kXcd, 0.0);
  size_t i = 0;
  while (i < tiles.size()) {
    size_t j = i + 1;
    while (j < tiles.size() && tiles[j].task == tiles[i].task && tiles[j].n0 == tiles[i].n0) ++j;
    const GemmTask& t = tasks[tiles[i].task];
    double k = 0.0;
    for (int q = 0; q < t.nseg; ++q) k += segs[t.seg0 + q].K;
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (size_t e = i; e < j; ++e) stream[x].push_back(tiles[e]);
    load[x] += (double)(j - i) * (k + 1.0);
    i = j;
  }
  std::vector<GemmTile> out;
  out.reserve(tiles.size());
  std::vector<size_t> pos(kXcd, 0);
  while (out.size() < tiles.size()) {
    for (int x = 0; x < kXcd && out.size() < tiles.size(); ++x) {
      int src = x;
      if (pos[src] >= stream[src].size()) {  // exhausted: borrow from the longest remaining stream
        size_t best = 0;
        for (int y = 0; y < kXcd; ++y)
          if (stream[y].size() - pos[y] > best) {
            best = stream[y].size() - pos[y];
            src = y;
          }
      }
      out.push_back(stream[src][pos[src]++]);
    }
  }
  return out;
}

GemmLaunch prepare_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs,
                        int cfg, int arith) {
  if (cfg < 0 || cfg >= kNumGemmShapes) throw Error(DMRJ_ERR_ARG, "GEMM tile config out of range");
  if (arith < 0 || arith > 2) throw Error(DMRJ_ERR_ARG, "GEMM arithmetic out of range");
  GemmLaunch g;
  g.cfg = cfg;
  g.arith = arith;
  const int bm = kGemmShapes[cfg].bm, bn = kGemmShapes[cfg].bn;
  // tiles ordered by descending K so that the persistent grid balances
  std::vector<std::pair<int64_t, int>> order;
  order.reserve(tasks.size());
  for (size_t i = 0; i < tasks.size(); ++i) {
    int64_t k = 0;
    for (int s = 0; s < tasks[i].nseg; ++s) k += segs[tasks[i].seg0 + s].K;
    order.push_back({-(k + 1), (int)i});
  }
  std::stable_sort(order.begin(), order.end());
  std::vector<GemmTile> tiles;
  for (auto& o : order) {
    const GemmTask& t = tasks[o.second];
    for (int n0 = 0; n0 < t.N; n0 += bn)
      for (int m0 = 0; m0 < t.M; m0 += bm) tiles.push_back({o.second, m0, n0, cfg});
  }
  if (tiles.empty()) return g;
  if (c.xcd_order) tiles = xcd_interleave(tiles, tasks, segs);
  UploadPack up;
  const size_t o_tiles = up.add(tiles.data(), tiles.size() * sizeof(GemmTile));
  const size_t o_tasks = up.add(tasks.data(), tasks.size() * sizeof(GemmTask));
  const size_t o_segs = up.add(segs.data(), segs.size() * sizeof(GemmSeg));
  g.ntiles = (int)tiles.size();
  g.nseg = (int)segs.size();
  // persistent grid: never more work-groups than can be resident at once
  const int64_t resident = (int64_t)c.num_cu * occupancy(c, cfg, arith);
  g.grid = (int)std::min<int64_t>((int64_t)tiles.size(), resident);
  if (arith == 2) {
    // per-segment operand maxima for the power-of-two scaling: work items
    // (segment, operand, first column, last column) of the stored operand
    std::vector<int4> mw;
    std::vector<int> st(std::max<size_t>(1, segs.size()), 0);
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      const GemmTask& t = tasks[ti];
      for (int q = 0; q < t.nseg; ++q) {
        const int si = t.seg0 + q;
        st[si] = (int)ti;
        const GemmSeg& sg = segs[si];
        if (sg.K <= 0 || t.M <= 0 || t.N <= 0) continue;
        const bool tA = sg.flags & 1, tB = sg.flags & 2;
        const int64_t ra = tA ? sg.K : t.M, ca = tA ? t.M : sg.K;
        const int64_t rb = tB ? t.N : sg.K, cb = tB ? sg.K : t.N;
        const int64_t stepa = std::max<int64_t>(1, 16384 / std::max<int64_t>(ra, 1));
        const int64_t stepb = std::max<int64_t>(1, 16384 / std::max<int64_t>(rb, 1));
        for (int64_t c0 = 0; c0 < ca; c0 += stepa) mw.push_back(make_int4(si, 0, (int)c0, (int)std::min(ca, c0 + stepa)));
        for (int64_t c0 = 0; c0 < cb; c0 += stepb) mw.push_back(make_int4(si, 1, (int)c0, (int)std::min(cb, c0 + stepb)));
      }
    }
    const size_t o_segtask = up.add(st.data(), st.size() * sizeof(int));
    g.segmax = (double*)c.alloc.alloc(std::max<size_t>(2, 2 * segs.size()) * sizeof(double));
    c.transient.push_back(g.segmax);  // lives until the next synchronisation point
    F16PlaneOffsets po;
    size_t o_maxwork = 0;
    if (c.f16_presplit) {
      po = prepare_f16_planes(c, g, tasks, segs, up);  // maxima per distinct operand block instead
    } else {
      o_maxwork = up.add(mw.data(), mw.size() * sizeof(int4));
      g.nmaxwork = (int)mw.size();
    }
    char* base = up.commit(c);
    g.segtask = (const int*)(base + o_segtask);
    if (g.presplit) {
      g.regionmaxwork = g.nregionmaxwork ? (const int4*)(base + po.regionmaxwork) : nullptr;
      g.regions = g.nregions ? (const F16Region*)(base + po.regions) : nullptr;
      g.splitwork = g.nsplitwork ? (const int4*)(base + po.splitwork) : nullptr;
      g.f16segs = (const F16Seg*)(base + po.f16segs);
    } else if (g.nmaxwork) {
      g.maxwork = (const int4*)(base + o_maxwork);
    }
    g.tiles = (const GemmTile*)(base + o_tiles);
    g.tasks = (const GemmTask*)(base + o_tasks);
    g.segs = (const GemmSeg*)(base + o_segs);
    return g;
  }
  char* base = up.commit(c);
  g.tiles = (const GemmTile*)(base + o_tiles);
  g.tasks = (const GemmTask*)(base + o_tasks);
  g.segs = (const GemmSeg*)(base + o_segs);
  return g;
}

static uint64_t histogram_key(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int arith) {
  double F = 0.0, lm = 0.0, ln = 0.0, lk = 0.0;
  for (const auto& t : tasks) {
    int64_t K = 0;
    for (int s = 0; s < t.nseg; ++s) K += segs[t.seg0 + s].K;
    const double f = (double)t.M * t.N * std::max<int64_t>(K, 1);
    F += f;
    lm += f * std::log2(std::max(t.M, 1));
    ln += f * std::log2(std::max(t.N, 1));
    lk += f * std::log2((double)std::max<int64_t>(K, 1));
  }
  auto bin = [](double v, int scale) { return (uint64_t)std::max(0, (int)std::lround(v * scale)) & 0xff; };
  uint64_t key = 0;
  if (F > 0) {
    key |= bin(lm / F, 2);
    key |= bin(ln / F, 2) << 8;
    key |= bin(lk / F, 2) << 16;
    key |= bin(std::log2(F), 1) << 24;
  }
  key |= bin(std::log2((double)tasks.size() + 1.0), 1) << 32;
  key |= (uint64_t)arith << 40;
  return key;
}

static bool repeatable(const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs) {
  struct Range {
    uintptr_t a, b;
  };
  std::vector<Range> in, out;
  for (const auto& t : tasks) {
    if (t.beta != 0.0) return false;
    if (t.M <= 0 || t.N <= 0) continue;
    out.push_back({(uintptr_t)t.C, (uintptr_t)(t.C + t.ldc * (t.N - 1) + t.M)});
    for (int s = 0; s < t.nseg; ++s) {
      const GemmSeg& g = segs[t.seg0 + s];
      if (g.K <= 0) continue;
      const bool tA = g.flags & 1, tB = g.flags & 2;
      const int64_t ea = tA ? g.lda * (t.M - 1) + g.K : g.lda * (g.K - 1) + t.M;
      const int64_t eb = tB ? g.ldb * (g.K - 1) + t.N : g.ldb * (t.N - 1) + g.K;
      in.push_back({(uintptr_t)g.A, (uintptr_t)(g.A + ea)});
      in.push_back({(uintptr_t)g.B, (uintptr_t)(g.B + eb)});
    }
  }
  auto by_start = [](const Range& x, const Range& y) { return x.a < y.a; };
  std::sort(in.begin(), in.end(), by_start);
  // merge the input ranges, then test every output range against them
  std::vector<Range> merged;
  for (const auto& r : in) {
    if (!merged.empty() && r.a <= merged.back().b)
      merged.back().b = std::max(merged.back().b, r.b);
    else
      merged.push_back(r);
  }
  for (const auto& o : out) {
    auto it = std::upper_bound(merged.begin(), merged.end(), Range{o.a, o.a}, by_start);
    if (it != merged.end() && it->a < o.b) return false;
    if (it != merged.begin() && std::prev(it)->b > o.a) return false;
  }
  return true;
}

int autotune_gemm(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int arith) {
  const uint64_t key = histogram_key(tasks, segs, arith);
  auto it = c.tune_cache.find(key);
  if (it != c.tune_cache.end()) {
    c.tune_hits += 1;
    return it->second;
  }
  if (c.autotune == 3) {  // frozen: keys not tuned yet take the heuristic, no trials
    c.tune_skipped += 1;
    return choose_gemm_cfg(tasks, c.num_cu);
  }
  if (!repeatable(tasks, segs)) {
    c.tune_skipped += 1;
    return choose_gemm_cfg(tasks, c.num_cu);
  }
  c.tune_misses += 1;
  constexpr int kReps = 2;
  hipEvent_t ev[kNumGemmShapes + 1];
  for (auto& e : ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing only
  GemmLaunch g[kNumGemmShapes];
  for (int cfg = 0; cfg < kNumGemmShapes; ++cfg) g[cfg] = prepare_gemm(c, tasks, segs, cfg, arith);
  auto issue = [&](const GemmLaunch& x) { issue_gemm(c, x); };
  launch_marker(c, 3);  // kernel-trace cut: tuning launches follow
  issue(g[0]);  // warm the operands into the caches once
  HIP_OK(hipEventRecord(ev[0], c.stream));
  for (int cfg = 0; cfg < kNumGemmShapes; ++cfg) {
    for (int r = 0; r < kReps; ++r) issue(g[cfg]);
    HIP_OK(hipEventRecord(ev[cfg + 1], c.stream));
  }
  launch_marker(c, 4);  // end of the tuning launches
  HIP_OK(hipEventSynchronize(ev[kNumGemmShapes]));
  int best = choose_gemm_cfg(tasks, c.num_cu);
  float best_ms = 1e30f, total = 0.f;
  for (int cfg = 0; cfg < kNumGemmShapes; ++cfg) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, ev[cfg], ev[cfg + 1]));
    total += ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  for (auto& e : ev) HIP_OK(hipEventDestroy(e));
  c.tune_ms += total;
  c.tune_cache[key] = best;
  return best;
}

int pick_gemm_cfg(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int arith) {
  if (c.gemm_cfg_override >= 0) return c.gemm_cfg_override;
  return c.autotune ? autotune_gemm(c, tasks, segs, arith) : choose_gemm_cfg(tasks, c.num_cu);
}

void issue_gemm(Context& c, const GemmLaunch& g) {
  if (g.arith == 1)
    issue_gemm_lp(c, g);
  else if (g.arith == 2)
    issue_gemm_f16(c, g);
  else
    issue_gemm_f64(c, g);
}

void launch_gemm_arith(Context& c, const std::vector<GemmTask>& tasks, const std::vector<GemmSeg>& segs, int cfg,
                       int arith) {
  if (tasks.empty()) return;
  if (cfg < 0) cfg = pick_gemm_cfg(c, tasks, segs, arith);
  if (arith == 0) c.last_gemm_cfg = cfg;
  GemmLaunch g = prepare_gemm(c, tasks, segs, cfg, arith);
  if (g.ntiles == 0) return;
  // one-kernel forms are timed by their own dispatch (hipExtLaunchKernelGGL
  // events); the FP16 form (scale pre-pass + GEMM) by events around both
  const int kind = arith == 1 ? KIND_GEMM_LP : (arith == 2 ? KIND_GEMM_F16 : KIND_GEMM_F64);
  if (arith == 2) {
    const int ps = c.prof_begin();
    issue_gemm(c, g);
    if (ps >= 0) c.prof_end(ps, kind, gemm_flops(tasks, segs), gemm_bytes(tasks, segs));
  } else {
    const int ps = c.prof_begin_ext();
    issue_gemm(c, g);
    c.prof_end_ext(ps, kind, gemm_flops(tasks, segs), gemm_bytes(tasks, segs));
  }
}

}  // namespace dmrj

// Pattern task plan (host, OpenMP).
#include "pattern_plan.hpp"

#include <omp.h>

#include <algorithm>

namespace mgpm {

namespace {

constexpr int kBuckets = 64;
constexpr uint64_t kGroupCost = uint64_t(1) << 16;
constexpr int32_t kGroupMax = 64;

inline int bucket_of(uint64_t cost) { return 63 - __builtin_clzll(cost | 1); }

// Orders items by descending log2(cost) (stable inside a bucket) and packs
// consecutive ones into groups of total cost <= kGroupCost.
template <class Item, class Cost>
void order_and_group(int64_t count, Cost cost_of, const std::vector<Item> &in,
                     std::vector<Item> &out, std::vector<int32_t> &groups,
                     std::vector<uint64_t> &group_cost) {
  std::vector<int64_t> hist(kBuckets, 0);
  std::vector<uint8_t> b(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; ++i) {
    b[i] = static_cast<uint8_t>(bucket_of(cost_of(i)));
    hist[b[i]]++;
  }
  std::vector<int64_t> start(kBuckets + 1, 0);
  for (int k = kBuckets - 1, acc = 0; k >= 0; --k) {
    start[k] = acc;
    acc += static_cast<int>(hist[k]);
  }
  out.resize(static_cast<size_t>(count));
  std::vector<uint64_t> cost(static_cast<size_t>(count));
  for (int64_t i = 0; i < count; ++i) {
    const int64_t pos = start[b[i]]++;
    out[pos] = in[i];
    cost[pos] = cost_of(i);
  }
  groups.assign(1, 0);
  group_cost.clear();
  uint64_t acc = 0;
  for (int64_t i = 0; i < count; ++i) {
    const int32_t open = static_cast<int32_t>(i) - groups.back();
    if (open > 0 && (acc + cost[i] > kGroupCost || open >= kGroupMax)) {
      groups.push_back(static_cast<int32_t>(i));
      group_cost.push_back(acc);
      acc = 0;
    }
    acc += cost[i];
  }
  if (groups.back() != count) {
    groups.push_back(static_cast<int32_t>(count));
    group_cost.push_back(acc);
  }
}

}  // namespace

namespace {

// Shares of a task of `work` estimated pair tests whose outer loops run
// over at most `outer` indices.
int32_t shares(uint64_t work, int64_t outer, uint64_t split_work) {
  const uint64_t want = (work + split_work - 1) / split_work;
  const int64_t cap =
      std::min<int64_t>(kPatternMaxSplit, std::max<int64_t>(1, outer));
  return static_cast<int32_t>(
      std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(want), cap)));
}

// Pair tests of pattern `p` on the task edge (v0, v1) with degrees d0, d1,
// bounded from the set algebra of pattern_kernels.cuh (L0/L1 are subsets
// of N(v0), N(v1) or their intersection, m = min(d0, d1)), and the length
// of the longest outer loop a share strides over.
struct EdgeCost {
  uint64_t work;
  int64_t outer;
};

EdgeCost edge_cost(int p, uint64_t d0, uint64_t d1) {
  const uint64_t m = std::min(d0, d1);
  const int64_t big = static_cast<int64_t>(std::max(d0, d1));
  switch (p) {
    case 1:  // pairs of I<(N0,N1) x D<(N0,N1); outer = the shorter side
      return {d0 * m + d0 + d1, static_cast<int64_t>(m)};
    case 2:  // per a in D(N0,N1): two counts; per z in D(N1,N0): an
             // intersection and the edges inside it
      return {(d0 + d1) * (d0 + 64), big};
    case 3:  // per v in D(N0,N1): an intersection and the edges above v
      return {d0 * (d0 + 64) + d1, static_cast<int64_t>(d0)};
    case 4:  // pairs of D(N0,N1) x D(N1,N0)
      return {d0 * d1 + d0 + d1, static_cast<int64_t>(m)};
    case 5:  // per element of I, D(N0,N1), D(N1,N0): counts; per c in I two
             // intersections and the edges between them
      return {(d0 + d1) * (m + 64), big};
    case 6:  // per c in I: two intersections and the edges between them;
             // per b in D(N1,N0): a count
      return {(m + d1) * (m + 64), static_cast<int64_t>(std::max(m, d1))};
    case 7:  // for y in I(N0,N1): scan it, pairs of its two parts
      return {m * (m * m / 4 + m) + d0 + d1, static_cast<int64_t>(m)};
    default:  // 8: as P2 on I(N0,N1)
      return {(m + d1) * (m + 64), static_cast<int64_t>(std::max(d1, m))};
  }
}

}  // namespace

void build_pattern_plan(const CsrGraph &g, int pattern, PatternPlan &plan,
                        uint64_t split_work) {
  if (split_work == 0) split_work = kPatternSplitWork;
  const int32_t n = g.n;
  plan.max_degree = g.max_degree();
  // P4 counts each undirected edge once, from its larger endpoint
  // (pattern_kernels.cuh:878-882 skips v1 >= v0); the others take all 2m
  // directed edges (graph.cc:336-365).
  auto keep = [&](int32_t v, int32_t u) { return pattern != 4 || u < v; };
  std::vector<int32_t> esplit(static_cast<size_t>(g.nnz), 0);
  std::vector<int64_t> eoff(static_cast<size_t>(n) + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; ++v) {
    int64_t acc = 0;
    for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e) {
      const int32_t u = g.colidx[e];
      if (!keep(v, u)) continue;
      const EdgeCost c = edge_cost(pattern, g.degree(v), g.degree(u));
      esplit[e] = shares(c.work + 1, c.outer, split_work);
      acc += esplit[e];
    }
    eoff[v + 1] = acc;
  }
  for (int32_t v = 0; v < n; ++v) eoff[v + 1] += eoff[v];
  std::vector<PatternTask> tasks(static_cast<size_t>(eoff[n]));
  std::vector<uint64_t> tcost(tasks.size());
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; ++v) {
    int64_t k = eoff[v];
    for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e) {
      const int32_t u = g.colidx[e];
      if (!keep(v, u)) continue;
      const EdgeCost c = edge_cost(pattern, g.degree(v), g.degree(u));
      // Shares of a split edge each get a queue entry of their own.
      const uint64_t w = c.work / uint64_t(esplit[e]) + 1;
      const uint64_t share = esplit[e] > 1 ? std::max(w, kGroupCost) : w;
      for (int32_t s = 0; s < esplit[e]; ++s, ++k) {
        tasks[k] = PatternTask{v, u, s, esplit[e]};
        tcost[k] = share;
      }
    }
  }
  order_and_group(
      static_cast<int64_t>(tasks.size()), [&](int64_t i) { return tcost[i]; },
      tasks, plan.edges, plan.edge_groups, plan.edge_group_cost);
  plan.verts.clear();
  plan.vert_groups.assign(1, 0);
  plan.vert_group_cost.clear();
  if (pattern != 1) return;
  // P1 vertices: d(d-1)/2 pair tests, outer loop over d.
  std::vector<PatternTask> vt;
  std::vector<uint64_t> vcost;
  vt.reserve(static_cast<size_t>(n));
  for (int32_t v = 0; v < n; ++v) {
    const uint64_t d = g.degree(v);
    const uint64_t work = (d + 1) * (d + 1) / 2 + 1;
    const int32_t ns = shares(work, static_cast<int64_t>(d), split_work);
    const uint64_t w = work / uint64_t(ns) + 1;
    const uint64_t share = ns > 1 ? std::max(w, kGroupCost) : w;
    for (int32_t s = 0; s < ns; ++s) {
      vt.push_back(PatternTask{v, 0, s, ns});
      vcost.push_back(share);
    }
  }
  order_and_group(
      static_cast<int64_t>(vt.size()), [&](int64_t i) { return vcost[i]; },
      vt, plan.verts, plan.vert_groups, plan.vert_group_cost);
}

}  // namespace mgpm

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

void build_pattern_plan(const CsrGraph &g, PatternPlan &plan) {
  const int32_t n = g.n;
  plan.max_degree = g.max_degree();
  std::vector<PatternTaskEdge> coo(static_cast<size_t>(g.nnz));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; ++v)
    for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e)
      coo[e] = PatternTaskEdge{v, g.colidx[e]};
  // Quadratic proxy of the per-edge set algebra: (d0 + 16) * (d1 + 16).
  order_and_group(
      g.nnz,
      [&](int64_t i) {
        return uint64_t(g.degree(coo[i].v0) + 16) *
               uint64_t(g.degree(coo[i].v1) + 16);
      },
      coo, plan.edges, plan.edge_groups, plan.edge_group_cost);
  std::vector<int32_t> ids(static_cast<size_t>(n));
  for (int32_t v = 0; v < n; ++v) ids[v] = v;
  order_and_group(
      n, [&](int64_t v) { return uint64_t(g.degree(v) + 16) * (g.degree(v) + 16); },
      ids, plan.verts, plan.vert_groups, plan.vert_group_cost);
}

}  // namespace mgpm

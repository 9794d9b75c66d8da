// Triangle-count plan builder (host, OpenMP).
#include "tc_plan.hpp"

#include <omp.h>

#include <algorithm>
#include <numeric>

#include "mgpm.h"

namespace mgpm {

int build_tc_plan(const CsrGraph &g, const TcPlanParams &p, TcPlan &plan) {
  Dag dag;
  build_dag(g, dag);
  const int32_t n = g.n;
  plan.n = n;
  plan.m = g.nnz / 2;
  plan.m_plus = dag.rowptr[n];
  plan.max_out = dag.max_out;
  if (plan.m_plus >= (int64_t(1) << 32)) return MGPM_ERR_CAPACITY;
  const std::vector<int32_t> &od = dag.out_deg;

  uint64_t splus = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : splus)
  for (int32_t u = 0; u < n; ++u)
    for (int64_t i = dag.rowptr[u]; i < dag.rowptr[u + 1]; ++i)
      splus += static_cast<uint64_t>(od[dag.colidx[i]]);
  plan.s_plus = splus;

  // Owner rule (graph.cc:253-263 applied to every edge): T owns {T,x} when
  // d+(T) > d+(x) or (equal and T > x); edges with an empty side carry no
  // triangle and are dropped (graph.cc:242).
  auto owns = [&](int32_t t, int32_t x) {
    const int32_t a = od[t], b = od[x];
    return a >= 1 && b >= 1 && (a > b || (a == b && t > x));
  };
  std::vector<int64_t> pc(n, 0), poff;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t t = 0; t < n; ++t) {
    int64_t c = 0;
    for (int64_t i = g.rowptr[t]; i < g.rowptr[t + 1]; ++i)
      c += owns(t, g.colidx[i]);
    pc[t] = c;
  }
  prefix_sum(pc, poff);
  plan.partners.resize(static_cast<size_t>(poff[n]));
  uint64_t probes = 0;
  std::vector<int64_t> nchunks(n, 0);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : probes)
  for (int32_t t = 0; t < n; ++t) {
    int64_t o = poff[t];
    int64_t chunks = 0, acc = 0;
    for (int64_t i = g.rowptr[t]; i < g.rowptr[t + 1]; ++i) {
      const int32_t x = g.colidx[i];
      if (!owns(t, x)) continue;
      TcPartner pr;
      pr.beg = static_cast<uint32_t>(dag.rowptr[x]);
      pr.len = static_cast<uint32_t>(od[x]);
      plan.partners[o++] = pr;
      probes += pr.len;
      const int64_t cost = int64_t(pr.len) + kTcPartnerCost;
      if (acc == 0 || acc + cost > p.chunk_work) {
        ++chunks;
        acc = 0;
      }
      acc += cost;
    }
    nchunks[t] = chunks;
  }
  plan.probes = probes;

  std::vector<int64_t> coff;
  prefix_sum(nchunks, coff);
  std::vector<TcTask> all(static_cast<size_t>(coff[n]));
  std::vector<int64_t> work(all.size());
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t t = 0; t < n; ++t) {
    if (!pc[t]) continue;
    int64_t k = coff[t];
    int64_t acc = 0;
    for (int64_t i = poff[t]; i < poff[t + 1]; ++i) {
      const int64_t cost = int64_t(plan.partners[i].len) + kTcPartnerCost;
      if (i == poff[t] || acc + cost > p.chunk_work) {
        if (i != poff[t]) ++k;
        all[k].tbeg = static_cast<uint32_t>(dag.rowptr[t]);
        all[k].tlen = od[t];
        all[k].pbeg = static_cast<uint32_t>(i);
        all[k].pcnt = 0;
        work[k] = kTcTableCost + od[t];
        acc = 0;
      }
      all[k].pcnt += 1;
      work[k] += cost;
      acc += cost;
    }
  }

  auto cls = [&](int32_t tlen) {
    if (tlen <= p.small_max) return kTcSmall;
    if (tlen <= p.wave_max) return kTcMedium;
    if (tlen <= p.block_max) return kTcBlock;
    return kTcGlobal;
  };
  std::vector<int64_t> order(all.size());
  std::iota(order.begin(), order.end(), 0);
  // Longest tasks first (LPT) so the dynamic queue drains evenly.
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (work[a] != work[b]) return work[a] > work[b];
    return a < b;
  });
  // Queue-entry size per class: at most group_work, and small enough that
  // the class yields about kTcTargetGroups entries (enough waves to fill
  // the device) -- small graphs otherwise give a few waves long serial
  // groups.
  int64_t class_work[kTcNumClasses] = {0, 0, 0, 0};
  for (size_t i = 0; i < all.size(); ++i) class_work[cls(all[i].tlen)] += work[i];
  int64_t limit[kTcNumClasses];
  for (int c = 0; c < kTcNumClasses; ++c)
    limit[c] = std::max<int64_t>(
        kTcMinGroupWork,
        std::min<int64_t>(p.group_work, class_work[c] / kTcTargetGroups));
  int64_t group_cost[kTcNumClasses] = {0, 0, 0, 0};
  for (int c = 0; c < kTcNumClasses; ++c) {
    plan.tasks[c].clear();
    plan.group_off[c].assign(1, 0);
    plan.group_cost[c].clear();
    plan.class_max_tlen[c] = 0;
  }
  for (int64_t i : order) {
    const int c = cls(all[i].tlen);
    // Close the open group of class c if this task would overflow it.
    std::vector<int32_t> &go = plan.group_off[c];
    const int32_t open = static_cast<int32_t>(plan.tasks[c].size()) - go.back();
    if (open > 0 && (group_cost[c] + work[i] > limit[c] || open >= 64)) {
      go.push_back(static_cast<int32_t>(plan.tasks[c].size()));
      plan.group_cost[c].push_back(group_cost[c]);
      group_cost[c] = 0;
    }
    plan.tasks[c].push_back(all[i]);
    group_cost[c] += work[i];
    plan.class_max_tlen[c] = std::max(plan.class_max_tlen[c], all[i].tlen);
  }
  for (int c = 0; c < kTcNumClasses; ++c)
    if (static_cast<int32_t>(plan.tasks[c].size()) != plan.group_off[c].back()) {
      plan.group_off[c].push_back(static_cast<int32_t>(plan.tasks[c].size()));
      plan.group_cost[c].push_back(group_cost[c]);
    }
  plan.dag_col.swap(dag.colidx);
  // The kernels gather partner and owner lists without bounds checks: every
  // list must lie inside the DAG column array.
  const uint64_t ncol = plan.dag_col.size();
  bool inside = true;
#pragma omp parallel for reduction(&& : inside)
  for (int64_t i = 0; i < static_cast<int64_t>(plan.partners.size()); ++i)
    inside = inside && uint64_t(plan.partners[i].beg) + plan.partners[i].len <= ncol;
  for (int c = 0; c < kTcNumClasses; ++c)
    for (const TcTask &t : plan.tasks[c])
      inside = inside && t.tlen >= 0 && uint64_t(t.tbeg) + uint64_t(t.tlen) <= ncol &&
               uint64_t(t.pbeg) + uint64_t(t.pcnt) <= plan.partners.size();
  if (!inside) return MGPM_ERR_INTERNAL;
  return MGPM_OK;
}

}  // namespace mgpm

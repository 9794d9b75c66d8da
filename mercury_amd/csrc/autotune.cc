// C ABI: autotuner for the counting kernels' knobs (no counterpart in the
// reference, whose schedule constants are fixed at compile time:
// tuning_schedules.cuh:41-47, tc_challenge.cu:67-73).
//
// Triangle counting: coordinate descent over the table-class thresholds
// (which owners get 256-slot wave tables, 2048-slot wave tables, workgroup
// tables or binary search), the task split and packing granularity, list
// items per lane and the streaming mode; every candidate is timed on the
// device (median of `reps` runs of the counting kernels) and must return
// the same count as the starting point.  Patterns: the threshold above
// which tasks are split into shares, and the persistent grid.  The winner stays
// resident on the graph, so later calls with a NULL tune, batches and the
// server all use it.
#include <algorithm>
#include <cstring>
#include <vector>

#include "capi_common.hpp"

using namespace mgpm;

namespace {

// TcPlanParams -> the public knob encoding (-1 = empty class).
mgpm_tune tune_of(const TcResident &r) {
  mgpm_tune t;
  std::memset(&t, 0, sizeof(t));
  const TcPlanParams &p = r.params;
  t.small_max_deg = p.small_max == 0 ? -1 : p.small_max;
  t.wave_max_deg = p.wave_max == p.small_max ? -1 : p.wave_max;
  t.block_max_deg = p.block_max == p.wave_max ? -1 : p.block_max;
  t.chunk_work = static_cast<int32_t>(p.chunk_work);
  t.group_work = static_cast<int32_t>(p.group_work);
  t.tc_items = r.items;
  t.grid_blocks = r.grid_cap;
  t.tc_cursor_min = r.cursor_min;
  t.tc_layout = r.layout;
  t.tc_merge_max = r.merge_max == 0 ? -1 : r.merge_max;
  t.tc_half_waves = r.half_waves;
  return t;
}

double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

int time_tc(mgpm_dgraph *dg, const mgpm_tune &t, int reps, uint64_t *count,
            double *ms) {
  std::vector<double> k;
  for (int i = 0; i < reps; ++i) {
    mgpm_stats st;
    const int rc = mgpm_tc_count(dg, &t, nullptr, count, &st);
    if (rc) return rc;
    if (i > 0 || reps == 1) k.push_back(st.kernel_ms);  // first builds the plan
  }
  *ms = median(k);
  return MGPM_OK;
}

int time_pattern(mgpm_dgraph *dg, int pattern, const mgpm_tune &t, int reps,
                 uint64_t *count, double *ms) {
  std::vector<double> k;
  for (int i = 0; i < reps; ++i) {
    mgpm_stats st;
    const int rc = mgpm_pattern_count(dg, pattern, &t, nullptr, count, &st);
    if (rc) return rc;
    k.push_back(st.kernel_ms);
  }
  *ms = median(k);
  return MGPM_OK;
}

// One knob of the descent: its field and candidate values.
struct Knob {
  int32_t mgpm_tune::*field;
  std::vector<int32_t> values;
};

int tune_tc(mgpm_dgraph *dg, int reps, mgpm_tune *best, double *best_ms) {
  mgpm_tune cur = tune_of(dg->tc);
  uint64_t want = 0;
  double cur_ms = 0;
  int rc = time_tc(dg, cur, reps, &want, &cur_ms);
  if (rc) return rc;
  const std::vector<Knob> knobs = {
      {&mgpm_tune::tc_layout, {-1, 0}},
      {&mgpm_tune::tc_merge_max, {-1, 16, 32, 64}},
      {&mgpm_tune::tc_half_waves, {-1, 1}},
      {&mgpm_tune::tc_items, {4, 8}},
      {&mgpm_tune::tc_cursor_min, {-1, 16, 32, 64}},
      {&mgpm_tune::small_max_deg, {32, 64, 128}},
      {&mgpm_tune::wave_max_deg, {256, 512, 1024}},
      {&mgpm_tune::block_max_deg, {-1, 4096, 16384}},
      {&mgpm_tune::chunk_work, {8192, 16384, 32768, 65536, 131072}},
      {&mgpm_tune::group_work, {8192, 32768, 131072}},
  };
  // A candidate must beat the incumbent by 2% to replace it.
  for (const Knob &k : knobs) {
    for (int32_t v : k.values) {
      if (cur.*k.field == v) continue;
      mgpm_tune cand = cur;
      cand.*k.field = v;
      uint64_t c = 0;
      double ms = 0;
      if ((rc = time_tc(dg, cand, reps, &c, &ms))) return rc;
      if (c != want)
        return fail(MGPM_ERR_INTERNAL, "autotune: candidate count differs");
      if (ms < 0.98 * cur_ms) {
        cur = cand;
        cur_ms = ms;
      }
    }
  }
  // Leave the winner resident (plan, items, grid cap).
  if ((rc = tc_replan(dg, tc_params_from(&cur, dg->host->nnz / 2), nullptr)))
    return rc;
  dg->tc.items = cur.tc_items;
  dg->tc.grid_cap = cur.grid_blocks;
  dg->tc.cursor_min = cur.tc_cursor_min;
  dg->tc.layout = cur.tc_layout;
  dg->tc.merge_max = cur.tc_merge_max > 0 ? cur.tc_merge_max : 0;
  dg->tc.half_waves = cur.tc_half_waves;
  dg->tc.tuned = true;
  *best = cur;
  *best_ms = cur_ms;
  return MGPM_OK;
}

int tune_pattern(mgpm_dgraph *dg, int pattern, int reps, mgpm_tune *best,
                 double *best_ms) {
  mgpm_tune cur;
  std::memset(&cur, 0, sizeof(cur));
  int rc = pattern_make_resident(dg, nullptr, pattern);
  if (rc) return rc;
  cur.grid_blocks = dg->pat.grid_cap;
  cur.pattern_split = dg->pat.plan[pattern].split;
  // The knobs tuned here steer the per-edge kernels (shares and grid); the
  // census ignores them, so every candidate is timed on those kernels.
  cur.pattern_census = -1;
  uint64_t want = 0;
  double cur_ms = 0;
  if ((rc = time_pattern(dg, pattern, cur, reps, &want, &cur_ms))) return rc;
  const std::vector<Knob> knobs = {
      {&mgpm_tune::pattern_split, {-1, 16, 18, 20, 22}},
      {&mgpm_tune::grid_blocks, {dg->cus, 2 * dg->cus, 4 * dg->cus}},
  };
  for (const Knob &k : knobs) {
    for (int32_t v : k.values) {
      if (cur.*k.field == v) continue;
      mgpm_tune cand = cur;
      cand.*k.field = v;
      uint64_t c = 0;
      double ms = 0;
      if ((rc = time_pattern(dg, pattern, cand, reps, &c, &ms))) return rc;
      if (c != want)
        return fail(MGPM_ERR_INTERNAL, "autotune: candidate count differs");
      if (ms < 0.98 * cur_ms) {
        cur = cand;
        cur_ms = ms;
      }
    }
  }
  if ((rc = pattern_make_resident(dg, nullptr, pattern, cur.pattern_split)))
    return rc;
  dg->pat.grid_cap = cur.grid_blocks;
  *best = cur;
  *best_ms = cur_ms;
  return MGPM_OK;
}

}  // namespace

namespace mgpm {

int autotune_size_classes(const std::vector<mgpm_dgraph *> &graphs, int reps) {
  auto cls = [](const mgpm_dgraph *g) {
    return 63 - __builtin_clzll(g->tc.m_plus | 1);
  };
  std::vector<mgpm_dgraph *> todo;
  for (mgpm_dgraph *g : graphs)
    if (!g->tc.tuned && std::find(todo.begin(), todo.end(), g) == todo.end())
      todo.push_back(g);
  std::sort(todo.begin(), todo.end(), [&](const mgpm_dgraph *a,
                                          const mgpm_dgraph *b) {
    return a->tc.m_plus > b->tc.m_plus;
  });
  for (size_t i = 0; i < todo.size();) {
    size_t j = i + 1;
    while (j < todo.size() && cls(todo[j]) == cls(todo[i])) ++j;
    mgpm_tune t;
    double ms = 0;
    int rc = tune_tc(todo[i], reps, &t, &ms);  // the class's largest graph
    if (rc) return rc;
    todo[i]->tc.generation = next_generation();
    const TcPlanParams p = tc_params_from(&t);
    for (size_t k = i + 1; k < j; ++k) {
      mgpm_dgraph *g = todo[k];
      if ((rc = tc_replan(g, p, nullptr))) return rc;
      g->tc.items = t.tc_items;
      g->tc.grid_cap = t.grid_blocks;
      g->tc.cursor_min = t.tc_cursor_min;
      g->tc.tuned = true;
    }
    i = j;
  }
  return MGPM_OK;
}

}  // namespace mgpm

extern "C" int mgpm_autotune(mgpm_dgraph *dg, int pattern, int32_t reps,
                             mgpm_tune *best, double *best_ms) {
  if (!dg) return fail(MGPM_ERR_INVALID, "null argument");
  if (pattern < MGPM_TC || pattern > MGPM_P8)
    return fail(MGPM_ERR_PATTERN, "Not support.");
  try {
    int rc = set_device(dg->device);
    if (rc) return rc;
    if (reps <= 0) reps = 3;
    mgpm_tune t;
    double ms = 0;
    rc = pattern == MGPM_TC ? tune_tc(dg, reps, &t, &ms)
                            : tune_pattern(dg, pattern, reps, &t, &ms);
    if (rc) return rc;
    dg->tc.generation = next_generation();  // batches keyed on this re-plan
    if (best) *best = t;
    if (best_ms) *best_ms = ms;
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::exception &ex) {
    return fail(MGPM_ERR_INTERNAL, ex.what());
  }
}

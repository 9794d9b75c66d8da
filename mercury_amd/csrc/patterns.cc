// C ABI: P1..P8 pattern counts (subgraph_matching/sm_multigpu.cu:97-239 on
// one device, without MPI).
#include <algorithm>
#include <cstring>

#include "capi_common.hpp"
#include "pattern_plan.hpp"

namespace mgpm {

namespace {

// Per-wave scratch is 8 * cap ints; the persistent grid is sized to fill
// the device but kept under this many bytes of scratch in total.
constexpr uint64_t kSlabBudget = uint64_t(8) << 30;

template <class T>
int upload(DevBuf &buf, const std::vector<T> &v, hipStream_t s) {
  MGPM_HIP(buf.alloc(v.size() * sizeof(T)));
  if (!v.empty())
    MGPM_HIP(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T),
                            hipMemcpyHostToDevice, s));
  return MGPM_OK;
}

}  // namespace

int pattern_make_resident(mgpm_dgraph *dg, hipStream_t s, int pattern,
                          int split) {
  PatternResident &r = dg->pat;
  PatternPlanResident &pr = r.plan[pattern];
  if (split == 0) split = pr.ready ? pr.split : r.split;
  if (pr.ready && split == pr.split) return MGPM_OK;
  const bool replace = pr.ready;
  PatternPlan plan;
  build_pattern_plan(*dg->host, pattern, plan,
                     split < 0 ? UINT64_MAX : uint64_t(1) << std::min(split, 62));
  // Upload into fresh buffers; install them only when every copy has
  // landed, so a failure leaves the resident plan untouched.
  DevBuf edges, edge_groups, verts, vert_groups;
  int rc;
  if ((rc = upload(edges, plan.edges, s))) return rc;
  if ((rc = upload(edge_groups, plan.edge_groups, s))) return rc;
  if ((rc = upload(verts, plan.verts, s))) return rc;
  if ((rc = upload(vert_groups, plan.vert_groups, s))) return rc;
  MGPM_HIP(hipStreamSynchronize(s));
  pr.edges.swap(edges);
  pr.edge_groups.swap(edge_groups);
  pr.verts.swap(verts);
  pr.vert_groups.swap(vert_groups);
  pr.n_edge_groups = static_cast<int>(plan.edge_groups.size()) - 1;
  pr.n_vert_groups = static_cast<int>(plan.vert_groups.size()) - 1;
  pr.edge_group_cost = std::move(plan.edge_group_cost);
  pr.vert_group_cost = std::move(plan.vert_group_cost);
  pr.split = split;
  pr.ready = true;
  r.cap = std::max<int64_t>(plan.max_degree, 64);
  if (replace) dg->tc.generation = next_generation();  // batches holding the old plan re-plan
  return MGPM_OK;
}

PatternDevicePlan pattern_device_plan(const mgpm_dgraph *dg, int pattern) {
  const PatternResident &r = dg->pat;
  const PatternPlanResident &pr = r.plan[pattern];
  PatternDevicePlan d;
  d.rowptr = dg->csr.rowptr.as<int64_t>();
  d.col = dg->csr.colidx.as<int32_t>();
  d.edges = pr.edges.as<int4>();
  d.edge_groups = pr.edge_groups.as<int32_t>();
  d.n_edge_groups = pr.n_edge_groups;
  d.verts = pr.verts.as<int4>();
  d.vert_groups = pr.vert_groups.as<int32_t>();
  d.n_vert_groups = pr.n_vert_groups;
  d.slab = r.slab.as<int32_t>();
  d.cap = r.cap;
  d.queues = reinterpret_cast<unsigned *>(dg->scratch.as<char>() + 128);
  d.counters =
      reinterpret_cast<unsigned long long *>(dg->scratch.as<char>() + 192);
  return d;
}

int pattern_make_slab(mgpm_dgraph *dg) {
  PatternResident &r = dg->pat;
  if (r.slab.p) return MGPM_OK;
  const int wpb = pattern_waves_per_block();
  int grid = dg->cus * 4;  // 4 workgroups of 4 waves per CU
  const uint64_t per_block = uint64_t(wpb) * 8 * r.cap * sizeof(int32_t);
  grid = static_cast<int>(
      std::max<uint64_t>(1, std::min<uint64_t>(grid, kSlabBudget / per_block)));
  MGPM_HIP(r.slab.alloc(per_block * grid));
  r.grid = grid;
  return MGPM_OK;
}

}  // namespace mgpm

using namespace mgpm;

extern "C" int mgpm_pattern_count(mgpm_dgraph *dg, int pattern,
                                  const mgpm_tune *tune, void *stream,
                                  uint64_t *count, mgpm_stats *stats) {
  if (!dg || !count) return fail(MGPM_ERR_INVALID, "null argument");
  if (pattern == MGPM_TC) return mgpm_tc_count(dg, tune, stream, count, stats);
  if (pattern < MGPM_P1 || pattern > MGPM_P8)
    return fail(MGPM_ERR_PATTERN, "Not support.");
  try {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = set_device(dg->device);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!(tune && tune->pattern_census < 0)) {
      float kms = 0;
      if ((rc = census_pattern_count(dg, pattern, s, count, &kms))) return rc;
      if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->kernel_ms = kms;
        stats->n = static_cast<uint64_t>(dg->host->n);
        stats->m = static_cast<uint64_t>(dg->host->nnz / 2);
        stats->path = 30 + pattern;
        stats->e2e_ms = ms_since(t0);
      }
      return MGPM_OK;
    }
    double prep = 0;
    const int split = tune ? tune->pattern_split : 0;
    const PatternPlanResident &pr = dg->pat.plan[pattern];
    if (!pr.ready || !dg->pat.slab.p || (split != 0 && split != pr.split)) {
      const auto tp = std::chrono::steady_clock::now();
      if ((rc = pattern_make_resident(dg, s, pattern, split))) return rc;
      if ((rc = pattern_make_slab(dg))) return rc;
      prep = ms_since(tp);
    }
    int grid = dg->pat.grid;
    const int cap = (tune && tune->grid_blocks > 0) ? tune->grid_blocks
                                                    : dg->pat.grid_cap;
    if (cap > 0) grid = std::min(grid, cap);
    MGPM_HIP(hipEventRecord(dg->ev0, s));
    if (pattern_device_launch(pattern_device_plan(dg, pattern), pattern, grid,
                              s))
      return fail(MGPM_ERR_HIP, "pattern kernel launch failed");
    MGPM_HIP(hipEventRecord(dg->ev1, s));
    MGPM_HIP(hipMemcpyAsync(dg->pinned, dg->scratch.as<char>() + 192, 16,
                            hipMemcpyDeviceToHost, s));
    MGPM_HIP(hipStreamSynchronize(s));
    float kms = 0;
    MGPM_HIP(hipEventElapsedTime(&kms, dg->ev0, dg->ev1));
    const uint64_t c1 = dg->pinned[0], c2 = dg->pinned[1];
    // sm_multigpu.cu:209-232 on one rank: P4, P6 report counters[1]; the
    // others counters[1] - counters[2] (P1: result1 - result0, the same).
    *count = (pattern == MGPM_P4 || pattern == MGPM_P6) ? c1 : c1 - c2;
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->kernel_ms = kms;
      stats->prep_ms = prep;
      stats->n = static_cast<uint64_t>(dg->host->n);
      stats->m = static_cast<uint64_t>(dg->host->nnz / 2);
      stats->path = 10 + pattern;
      stats->e2e_ms = ms_since(t0);
    }
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::exception &ex) {
    return fail(MGPM_ERR_INTERNAL, ex.what());
  }
}

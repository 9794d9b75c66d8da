// Internal definitions shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "host_graph.hpp"
#include "mgpm.h"
#include "pattern_device.hpp"
#include "tc_device.hpp"
#include "tc_plan.hpp"

struct mgpm_graph {
  std::shared_ptr<mgpm::CsrGraph> g;
};

namespace mgpm {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define MGPM_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::mgpm::hip_fail(e_, #call); \
  } while (0)

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now() - t0)
      .count();
}

// Device buffer owned by RAII.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    reset();
    if (n == 0) n = 64;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    else p = nullptr;
    return e;
  }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Triangle-count state resident on one device.
struct TcResident {
  TcPlanParams params;
  uint64_t generation = 0;  // renewed (next_generation) whenever any plan
                            // of the graph (TC or pattern) or its launch
                            // settings change
  int items = 8;            // list items per lane per step
  int grid_cap = 0;         // cap on persistent workgroups (0 = none)
  int cursor_min = -1;      // owner-cursor streaming threshold (<= 0: off)
  bool tuned = false;       // knobs chosen by the autotuner
  bool on_device = false;   // plan built by tc_make_resident_device
  uint64_t s_plus = 0, probes = 0, m_plus = 0;
  uint64_t npart = 0;       // partners of the plan
  int32_t max_out = 0;
  int layout = 0;           // tc_layout choice of the resident plan
  int merge_max = 0;        // tc_merge_max choice (0: never)
  int half_waves = 0;       // tc_half_waves choice (0: the default)
  DevBuf col, partners, tasks[kTcNumClasses], groups[kTcNumClasses];
  // The plan's partner lists laid out task by task (tc_layout), built on
  // first use and dropped with the plan.
  DevBuf laid, laid_off;
  TcDevicePlan dev;
  size_t bytes() const;
  void drop_layout() {
    laid.reset();
    laid_off.reset();
  }
};

// Undirected CSR resident on one device (pattern kernels).
struct CsrResident {
  DevBuf rowptr, colidx;
  size_t bytes() const { return rowptr.bytes + colidx.bytes; }
};

// One pattern's task plan on the device, built on first use.
struct PatternPlanResident {
  bool ready = false;
  int split = 0;  // shares above 2^split pair tests (-1: no splits)
  int n_edge_groups = 0, n_vert_groups = 0;
  DevBuf edges, edge_groups, verts, vert_groups;
  std::vector<uint64_t> edge_group_cost, vert_group_cost;  // host copies
  size_t bytes() const {
    return edges.bytes + edge_groups.bytes + verts.bytes + vert_groups.bytes;
  }
};

// P1..P8 plans (one per pattern) and the per-wave scratch of the
// single-query path, allocated on its first call.
struct PatternResident {
  int grid = 0;      // workgroups the single-query slab holds
  int grid_cap = 0;  // cap on persistent workgroups (0 = none)
  int split = 20;    // split exponent for plans built from now on
  int64_t cap = 64;  // slab ints per list (>= max degree)
  DevBuf slab;
  PatternPlanResident plan[9];  // index = pattern (1..8)
  size_t bytes() const {
    size_t b = slab.bytes;
    for (const auto &p : plan) b += p.bytes();
    return b;
  }
};

// Device-built DAG and small/large division (mgpm_dgraph_orientation).
struct DagResident {
  bool ready = false;
  int32_t deg_thd = 0;
  int64_t sizes[3] = {0, 0, 0};  // DAG, small, large arcs
  DevBuf rowptr[3], col[3];       // outputs
  DevBuf deg, nd, src, val, pos, scan;  // work arrays
  size_t bytes() const {
    size_t b = deg.bytes + nd.bytes + src.bytes + val.bytes + pos.bytes +
               scan.bytes;
    for (int k = 0; k < 3; ++k) b += rowptr[k].bytes + col[k].bytes;
    return b;
  }
};

// The degree-oriented DAG the census counts over, built on the device on
// the census's first call and kept with the graph, with the pair-table
// layout of the 5-vertex census (a function of the graph alone).
struct CensusGraph {
  bool ready = false;
  int32_t max_deg = 1;   // largest degree
  int64_t max_out = 1;   // largest DAG out-degree
  // The graph relabelled in DAG order (census_relabel: ids are ranks by
  // degree, then id), its CSR and DAG; every census pass runs on it.
  DevBuf rowptr, col;
  DevBuf deg, src, drp, dcol, dsrc;
  bool pairs_ready = false;
  uint64_t tab_total = 0;  // pair-table slots
  DevBuf tab_off;          // n + 1 slot offsets
  size_t bytes() const {
    return rowptr.bytes + col.bytes + deg.bytes + src.bytes + drp.bytes +
           dcol.bytes + dsrc.bytes + tab_off.bytes;
  }
};

// Census workspace of one caller (the single-query path of a device, a
// batch lane, a server slot): grows to the largest query it has served and
// is reused without allocating.
struct CensusWork {
  // Bytes the pair passes' per-workgroup arrays may take (more workgroups
  // in flight on larger budgets): the single-query workspace takes 16 GiB,
  // batch lanes and server slots less.
  size_t budget = size_t(16) << 30;
  DevBuf t, t_csr, tv, small, slab, wedges, touched, k5_slab;
  DevBuf stamp, lo, mask, list, grows, anchors, table;
  size_t bytes() const {
    return t.bytes + t_csr.bytes + tv.bytes + small.bytes + slab.bytes +
           wedges.bytes + touched.bytes + k5_slab.bytes + stamp.bytes +
           lo.bytes + mask.bytes + list.bytes + grows.bytes + anchors.bytes +
           table.bytes;
  }
};

// Default task-cost chunk for a DAG of `arcs` arcs: 64K below 2^27 arcs,
// else 64K times the largest power of two <= arcs / 2^26, at most 1M.  On
// one MI355X, RMAT-22 and ER-22 (just under 2^26 arcs) count fastest with
// 64K chunks (18.11 vs 18.52 ms at 128K on RMAT-22), RMAT-24 (2.6 x 10^8
// arcs) with 128K (132 vs 175 ms): the workgroup-table owners of larger
// graphs rebuild their tables once per chunk.
inline int64_t tc_default_chunk(int64_t arcs) {
  int64_t f = 1;
  while (f < 16 && (arcs >> 26) >= 2 * f) f *= 2;
  return int64_t(65536) * f;
}

// A value no plan generation of any graph in this process has taken yet.
// Batches key their recorded plans on (graph, pattern, generation); a
// process-wide sequence keeps a graph uploaded at the address of a freed
// one from matching the freed graph's key.
uint64_t next_generation();

// Shared by the C-ABI units.  With arcs > 0 the unset chunk size takes
// tc_default_chunk(arcs).
TcPlanParams tc_params_from(const mgpm_tune *t, int64_t arcs = 0);
int set_device(int device);
int tc_make_resident(mgpm_dgraph *dg, const TcPlanParams &params,
                     hipStream_t s);
// The same plan built on the device from the resident CSR (tc_dplan.cc).
int tc_make_resident_device(mgpm_dgraph *dg, const TcPlanParams &params,
                            hipStream_t s);
// Items per lane the counting kernels use when no tune says otherwise:
// short streamed lists on average (fewer than 32 items per partner, e.g.
// uniform random graphs) waste most of an 8-row step, so they take 4.
inline int default_tc_items(uint64_t probes, uint64_t partners) {
  return probes < 32 * partners ? 4 : 8;
}
// P1..P8 in closed form from a subgraph census (census.cc): triangles,
// 4-cycles, 4- and 5-cliques for P1, P4, P7 and P8, and the 5-vertex pair
// passes for P2, P3, P5 and P6.
int census_pattern_count(mgpm_dgraph *dg, int pattern, hipStream_t s,
                         uint64_t *count, float *kernel_ms);
// The census split for callers that run many queries.  A census runs the
// triangle pass and its sums plus the passes of a census_passes mask; one
// run serves every pattern whose passes it includes (census_passes of P2,
// P3, P5 and P6 includes every pass, so one run serves all eight).
// census_prepare builds what the graph keeps (DAG, pair-table layout; may
// allocate and synchronise `s`) and grows `w` for the passes;
// census_enqueue then enqueues the whole census on `s` with no allocation
// and no host synchronisation, leaving the accumulators at census_acc(w);
// census_combine turns them into a pattern's count (MGPM_ERR_INTERNAL if
// the census's self-checks disagree).
enum CensusPass : int {
  kCensusPassC4 = 1,     // 4-cycles (P4)
  kCensusPassK5 = 2,     // 4-cliques per triangle, 5-cliques (P7, P8)
  kCensusPassPairs = 4,  // the 5-vertex pair passes (P2, P3, P5, P6)
};
int census_passes(int pattern);
int census_prepare(mgpm_dgraph *dg, int passes, CensusWork &w, hipStream_t s);
// What census_prepare builds for the graph alone (DAG and pair-table
// layout), with temporary buffers only.
int census_graph_prepare(mgpm_dgraph *dg, hipStream_t s);
int census_enqueue(mgpm_dgraph *dg, int passes, CensusWork &w, hipStream_t s);
unsigned long long *census_acc(const CensusWork &w);
int census_combine(int pattern, const unsigned long long *acc,
                   uint64_t *count);
// Device bytes of a census workspace for these passes on this graph
// (before the graph's census DAG exists: a bound that includes it).
uint64_t census_work_bytes(const mgpm_dgraph *dg, int passes, size_t budget);
// Rebuilds the resident triangle plan where it was built (host or device).
int tc_replan(mgpm_dgraph *dg, const TcPlanParams &params, hipStream_t s);
// Builds pattern `pattern`'s plan if missing or if `split` (0 = the
// graph's current setting) differs from the resident plan's.
int pattern_make_resident(mgpm_dgraph *dg, hipStream_t s, int pattern,
                          int split = 0);
// Device view of pattern `pattern`'s plan with the graph's own slab,
// queues and counters.
PatternDevicePlan pattern_device_plan(const mgpm_dgraph *dg, int pattern);
int pattern_make_slab(mgpm_dgraph *dg);
// Tunes the triangle knobs of the largest not-yet-tuned graph of every
// size class (log2 of DAG arcs) among `graphs` and applies the winner to
// the others of its class.
int autotune_size_classes(const std::vector<mgpm_dgraph *> &graphs, int reps);

}  // namespace mgpm

struct mgpm_dgraph {
  int device = 0;
  int cus = 0;
  std::shared_ptr<mgpm::CsrGraph> host;
  mgpm::TcResident tc;
  mgpm::CsrResident csr;
  mgpm::PatternResident pat;
  mgpm::DagResident dag;
  mgpm::CensusGraph census;            // DAG of the census
  mgpm::DevBuf scratch;               // queues + result counters
  unsigned long long *pinned = nullptr;  // host mirror of results
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

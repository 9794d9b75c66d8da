// Multi-query launcher shared by mgpm_batch and the server's slots.
//
// A set of (graph, pattern) queries runs on the same counting kernels as a
// single mgpm_tc_count / mgpm_pattern_count call, without a host round trip
// per query: one fill kernel zeroes every query's queues and counters, the
// queries are spread over `lanes` streams (longest-processing-time first,
// each to the least-loaded lane), every query's kernels are capped at the
// workgroups its queue can feed so small queries leave room for their
// neighbours on the other lanes, and one copy brings all counters back.
// Each query owns its queues and counters; pattern scratch is per lane, so
// queries on one lane run in order and queries on different lanes overlap.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "capi_common.hpp"

namespace mgpm {

class BatchExec {
 public:
  BatchExec() = default;
  BatchExec(const BatchExec &) = delete;
  BatchExec &operator=(const BatchExec &) = delete;
  ~BatchExec();

  // Creates the lane streams and events on `device` (current device).
  // census: P1..P8 in closed form from the census (one census workspace
  // per lane), else by the per-edge kernels (one slab per lane).
  int init(int device, int cus, int lanes, bool census = true);

  // Validates the queries, builds missing pattern plans (or the graphs'
  // census DAGs), assigns queries to lanes and sizes the workspaces.  Fails
  // with MGPM_ERR_CAPACITY, before allocating, when the workspaces would
  // grow past `max_bytes` in total.
  int plan(const mgpm_query *q, int nq, uint64_t max_bytes);

  // Device bytes plan() may add to the workspaces for these queries (a
  // bound: every lane grown to the largest query's needs; 0 if they fit
  // what is there).  query_need folds one query into the running maxima
  // that growth_for turns into that bound, for callers growing a set.
  uint64_t growth(const mgpm_query *q, int nq) const;
  void query_need(const mgpm_query &q, uint64_t *want_census,
                  uint64_t *want_slab) const;
  uint64_t growth_for(uint64_t want_census, uint64_t want_slab, int nq) const;

  // Frees the lanes' pattern workspaces (census and slabs) after the lanes
  // drain; the next plan() grows them again.
  void release();

  // Enqueues the planned queries on `s`: zero fill, fork to the lanes,
  // kernels, join back to `s`, copy of the counters to host memory.  The
  // counters are valid once `s` has reached this point.
  int enqueue(hipStream_t s, bool copy_back = true);

  // Enqueues only the copy of the counters to host memory on `s`.
  int enqueue_copy(hipStream_t s);

  // Count of planned query i (after the copy has completed); returns
  // MGPM_OK, or MGPM_ERR_INTERNAL if a census self-check failed.
  int result(int i, uint64_t *count) const;

  int size() const { return static_cast<int>(queries_.size()); }
  int lanes() const { return static_cast<int>(lanes_.size()); }
  uint64_t footprint() const;

 private:
  struct Planned {
    mgpm_dgraph *g = nullptr;
    int pattern = 0;
    uint64_t cost = 0;
    int job = -1;  // census job that gives its count (-1: own kernels)
  };
  // One census run per graph in the batch, with the union of the passes its
  // queries need.
  struct Job {
    mgpm_dgraph *g = nullptr;
    int passes = 0;
    uint64_t cost = 0;
  };
  bool census_ = true;
  std::vector<Job> jobs_;
  std::unique_ptr<CensusWork[]> census_work_;  // per lane
  int device_ = 0;
  int cus_ = 0;
  std::vector<hipStream_t> lanes_;
  std::vector<hipEvent_t> lane_done_;
  hipEvent_t fork_ = nullptr;
  std::unique_ptr<DevBuf[]> slab_;   // per lane
  std::vector<int64_t> slab_cap_;    // ints per list the slab is laid out for
  std::vector<int> slab_grid_;       // workgroups the slab holds
  std::vector<std::vector<int>> order_;  // per lane, in order: query i, or
                                         // census job j as -1 - j
  std::vector<Planned> queries_;
  DevBuf scratch_;                   // per query: 2 u64 counters + 4 u32
                                     // queues, then the census accumulators
  unsigned long long *pinned_ = nullptr;
  size_t pinned_n_ = 0;              // queries the pinned buffer holds
};

// Estimated device work of one query, in the units the plans count
// (TC: probes plus DAG arcs; patterns: the pattern's plan task costs).
uint64_t query_cost(const mgpm_dgraph *g, int pattern);

}  // namespace mgpm

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
  int init(int device, int cus, int lanes);

  // Validates the queries, builds missing pattern plans, assigns queries to
  // lanes and sizes the workspaces.  Fails with MGPM_ERR_CAPACITY, before
  // allocating, when the workspaces would grow past `max_bytes` in total.
  int plan(const mgpm_query *q, int nq, uint64_t max_bytes);

  // Enqueues the planned queries on `s`: zero fill, fork to the lanes,
  // kernels, join back to `s`, copy of the counters to host memory.  The
  // counters are valid once `s` has reached this point.
  int enqueue(hipStream_t s, bool copy_back = true);

  // Enqueues only the copy of the counters to host memory on `s`.
  int enqueue_copy(hipStream_t s);

  // Combined count of planned query i (after the copy has completed).
  uint64_t count(int i) const;

  int size() const { return static_cast<int>(queries_.size()); }
  int lanes() const { return static_cast<int>(lanes_.size()); }
  uint64_t footprint() const;

 private:
  struct Planned {
    mgpm_dgraph *g = nullptr;
    int pattern = 0;
    uint64_t cost = 0;
  };
  int device_ = 0;
  int cus_ = 0;
  std::vector<hipStream_t> lanes_;
  std::vector<hipEvent_t> lane_done_;
  hipEvent_t fork_ = nullptr;
  std::unique_ptr<DevBuf[]> slab_;   // per lane
  std::vector<int64_t> slab_cap_;    // ints per list the slab is laid out for
  std::vector<int> slab_grid_;       // workgroups the slab holds
  std::vector<std::vector<int>> order_;  // per lane, query indices in order
  std::vector<Planned> queries_;
  DevBuf scratch_;                   // 2 u64 counters + 4 u32 queues each
  unsigned long long *pinned_ = nullptr;
  size_t pinned_n_ = 0;              // queries the pinned buffer holds
};

// Estimated device work of one query, in the units the plans count
// (TC: probes plus DAG arcs; patterns: the pattern's plan task costs).
uint64_t query_cost(const mgpm_dgraph *g, int pattern);

}  // namespace mgpm

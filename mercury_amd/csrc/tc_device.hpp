// Device-side triangle-count plan and its launcher.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "tc_plan.hpp"

namespace mgpm {

// Capacities the kernels are compiled for (tables keep load factor <= 1/2).
constexpr int kTcSmallMaxDeg = 128;     // 256-slot wave tables
constexpr int kTcWaveMaxDeg = 1024;     // 2048-slot wave tables
constexpr int kTcBlockLogpMax = 15;     // 32768-slot workgroup tables
constexpr int kTcBlockMaxDeg = 16384;

struct TcDevicePlan {
  const int32_t *col = nullptr;       // DAG column array
  uint32_t ncol = 0;                  // its length (m+)
  const TcPartner *partners = nullptr;
  // Optional layout of the partner lists task by task (tc_layout): when
  // both are set the kernels stream items[item_off[t.pbeg] ..
  // item_off[t.pbeg + t.pcnt]) for task t instead of gathering its lists.
  const int32_t *items = nullptr;
  const uint64_t *item_off = nullptr;
  const TcTask *tasks[kTcNumClasses] = {nullptr, nullptr, nullptr, nullptr};
  int ntasks[kTcNumClasses] = {0, 0, 0, 0};
  const int32_t *groups[kTcNumClasses] = {nullptr, nullptr, nullptr, nullptr};
  int ngroups[kTcNumClasses] = {0, 0, 0, 0};
  int class_max_tlen[kTcNumClasses] = {0, 0, 0, 0};
  unsigned *queues = nullptr;         // kTcNumClasses counters (scratch)
  unsigned long long *result = nullptr;
};

struct TcLaunch {
  int cus = 256;      // compute units of the device
  int grid_cap = 0;   // cap on persistent workgroups per kernel (0 = none)
  int items = 8;      // list items per lane per step (4 or 8)
  bool reset = true;  // zero queues and result first (else the caller has)
  int cursor_min = -1;  // owner-cursor streaming from this average list
                        // length of a 64-partner batch (<= 0: never)
  int merge_max = 0;    // owners of d+ up to this (at most 64) merge their
                        // partner lists lane by lane instead of hashing
  bool half_small = false;  // laid-out lists: small-class tasks two per
                            // wave (tc_half_tables)
};

// Enqueues reset + counting kernels on `stream`; the count lands in
// *p.result.  Returns 0, or -1 if a launch failed.
int tc_device_launch(const TcDevicePlan &p, const TcLaunch &l,
                     hipStream_t stream);

// Enqueues a fill of `words` 32-bit zeros at `p` (a kernel, so it may be
// recorded into a HIP graph).  Returns 0, or -1 if the launch failed.
int device_zero(void *p, size_t words, hipStream_t stream);

// One-time per-process kernel attributes (large dynamic LDS).
int tc_device_init();

// The partner lists laid out in partner order, so that each task's items
// are contiguous: item_off[0..np] = exclusive scan of partners[].len (lens,
// np + 1 words, is scratch), items[item_off[i] .. item_off[i + 1]) = the
// list of partner i.  `items` holds item_off[np] words; scan_tmp holds
// tc_layout_scan_bytes(np) bytes.  Returns 0, or -1 if a launch failed.
size_t tc_layout_scan_bytes(uint64_t np);
int tc_layout(const int32_t *col, const TcPartner *partners, uint64_t np,
              uint64_t *lens, uint64_t *item_off, int32_t *items,
              void *scan_tmp, size_t scan_bytes, int cus, hipStream_t stream);

}  // namespace mgpm

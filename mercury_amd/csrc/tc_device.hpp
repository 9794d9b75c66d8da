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

}  // namespace mgpm

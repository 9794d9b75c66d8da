// Device-side P1..P8 plan and launcher.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace mgpm {

struct PatternDevicePlan {
  const int64_t *rowptr = nullptr;  // undirected CSR
  const int32_t *col = nullptr;
  const int4 *edges = nullptr;      // task-edge shares (v0, v1, sub, nsub),
                                    // cost-ordered
  const int32_t *edge_groups = nullptr;
  int n_edge_groups = 0;
  const int4 *verts = nullptr;      // P1 vertex shares (v, -, sub, nsub)
  const int32_t *vert_groups = nullptr;
  int n_vert_groups = 0;
  int32_t *slab = nullptr;          // per-wave scratch, 8 * cap ints each
  int64_t cap = 0;                  // >= max degree
  unsigned *queues = nullptr;       // 2 counters (edges, vertices)
  unsigned long long *counters = nullptr;  // c1, c2
};

int pattern_waves_per_block();

// Enqueues the kernels of `pattern` (1..8) on `s` with at most `grid`
// workgroups (the slab must hold grid * waves-per-block waves), preceded by
// a reset of queues and counters unless the caller has zeroed them.
int pattern_device_launch(const PatternDevicePlan &p, int pattern, int grid,
                          hipStream_t s, bool reset = true);

}  // namespace mgpm

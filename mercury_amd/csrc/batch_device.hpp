// Device-side description of a batch: many (graph, pattern) queries served
// by one persistent launch over a shared, cost-ordered work queue.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "tc_plan.hpp"

namespace mgpm {

struct BatchQuery {
  int32_t pattern;  // 0 = TC, 1..8 = P1..P8
  uint32_t tc_ncol;
  const int32_t *tc_col;
  const TcPartner *tc_partners;
  const TcTask *tc_tasks[kTcNumClasses];
  const int32_t *tc_groups[kTcNumClasses];
  const int64_t *rowptr;  // undirected CSR (patterns)
  const int32_t *col;
  const int2 *edges;
  const int32_t *edge_groups;
  const int32_t *verts;
  const int32_t *vert_groups;
};

enum BatchItemKind : int32_t {
  kItemTc = 0,       // TC task group `group` of class `cls`
  kItemEdges = 1,    // pattern task-edge group `group`
  kItemVerts = 2,    // P1 vertex group `group`
};

struct BatchItem {
  int32_t query;
  int32_t kind;
  int32_t group;
  int32_t cls;
};

struct BatchDevice {
  const BatchQuery *queries = nullptr;
  const BatchItem *items = nullptr;
  int nitems = 0;
  int nqueries = 0;
  unsigned *queue = nullptr;               // 1 counter
  unsigned long long *counters = nullptr;  // 2 per query
  int32_t *slab = nullptr;                 // pattern scratch per wave
  int64_t cap = 64;
};

// Enqueues reset + the batch kernel (no memset nodes, so it can be
// captured into a HIP graph).
int batch_device_launch(const BatchDevice &b, int grid, hipStream_t s);
int batch_waves_per_block();

}  // namespace mgpm

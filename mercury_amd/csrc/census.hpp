// Subgraph census by triangle enumeration (census_kernels.hip): per-edge
// triangle counts and the sums from which P1 (induced claws) follows in
// closed form.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace mgpm {

// Accumulator slots (uint64, zeroed by the caller).
enum CensusAcc : int {
  kCensusTriangles = 0,   // triangles
  kCensusTriDegrees = 1,  // sum over triangles of d(x) + d(y) + d(z) - 6
  kCensusK4 = 2,          // 4-cliques
  kCensusDeg3 = 3,        // sum over vertices of C(d, 3)
  kCensusEdgePairs = 4,   // sum over edges of C(t_e, 2)
  kCensusPathPairs = 5,   // sum over edges {x,y} of (d_x - t - 1)(d_y - t - 1)
  kCensusC4 = 6,          // 4-cycles (as subgraphs, not necessarily induced)
  kCensusSlots = 8
};

struct CensusArgs {
  const int64_t *drp = nullptr;  // DAG row offsets (n + 1)
  const int32_t *dcol = nullptr; // DAG column array (m)
  int32_t *dsrc = nullptr;       // work: source of every DAG arc (m)
  const int32_t *deg = nullptr;  // undirected degrees (n)
  int32_t n = 0;
  int64_t m = 0;
  uint32_t *t = nullptr;         // per DAG arc: triangles on the edge (zeroed)
  int32_t *slab = nullptr;       // census_waves() x slab_len ints
  int64_t slab_len = 0;          // >= the largest DAG out-degree
  unsigned *cursor = nullptr;    // arc cursor, then vertex cursor (zeroed)
  unsigned long long *acc = nullptr;  // kCensusSlots (zeroed)
  // 4-cycle pass (census_c4_launch): undirected CSR, and per workgroup a
  // dense counter array (zeroed; left zeroed) and a list, n entries each.
  const int64_t *rowptr = nullptr;
  const int32_t *col = nullptr;
  uint32_t *wedges = nullptr;
  int32_t *touched = nullptr;
  int c4_groups = 0;
};

// Waves of the enumeration kernel (slab rows to allocate).
int census_waves(int cus);

// Enqueues the arc sources, the enumeration and the sums.  Returns 0, or
// -1 if a launch failed.
int census_launch(const CensusArgs &c, int cus, hipStream_t s);

// Enqueues the 4-cycle count into acc[kCensusC4], with c4_groups
// workgroups; cursor[1] must be zero.  Returns 0, or -1 if a launch failed.
int census_c4_launch(const CensusArgs &c, hipStream_t s);

}  // namespace mgpm

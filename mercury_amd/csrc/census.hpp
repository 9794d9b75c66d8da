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
  kCensusEdgeTriples = 7, // sum over edges of C(t_e, 3)
  // census_k5_launch, with k_T = 4-cliques on triangle T:
  kCensusTriPairs = 8,    // sum over triangles of C(k_T, 2)
  kCensusTriEdgeK = 9,    // sum over triangles of k_T (t_ab + t_ac + t_bc - 6)
  kCensusTriDegK = 10,    // sum over triangles of k_T (d_a + d_b + d_c - 9)
  kCensusK4Rows = 11,     // 4-cliques again (from the adjacency rows)
  kCensusK5 = 12,         // 5-cliques
  kCensusMismatch = 13,   // arcs whose common neighbourhood != t_e (0)
  kCensusSlots = 16
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
  // 5-clique pass (census_k5_launch): per wave k5_len ints of slab
  // (census_k5_slab_len), k5_waves waves; max_deg >= every degree.
  int32_t *k5_slab = nullptr;
  int64_t k5_len = 0;
  int32_t max_deg = 0;
  int k5_waves = 0;
};

// Waves of the enumeration kernel (slab rows to allocate).
int census_waves(int cus);

// Enqueues the arc sources, the enumeration and the sums.  Returns 0, or
// -1 if a launch failed.
int census_launch(const CensusArgs &c, int cus, hipStream_t s);

// Per-wave slab ints of the 5-clique pass for DAG out-degrees up to
// max_out and degrees up to max_deg.
int64_t census_k5_slab_len(int64_t max_out, int64_t max_deg);

// Enqueues the per-triangle 4-clique terms and the 5-clique count
// (kCensusTriPairs .. kCensusMismatch), with k5_waves waves; needs t[] from
// census_launch and cursor[2] zero.  Returns 0, or -1 if a launch failed.
int census_k5_launch(const CensusArgs &c, hipStream_t s);

// Enqueues the 4-cycle count into acc[kCensusC4], with c4_groups
// workgroups; cursor[1] must be zero.  Returns 0, or -1 if a launch failed.
int census_c4_launch(const CensusArgs &c, hipStream_t s);

}  // namespace mgpm

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
  // Subgraph counts of the 5-vertex census (P2, P3, P5, P6; full5 set):
  kCensusDeg4 = 14,       // sum over vertices of C(d, 4)
  kCensusFork = 15,       // sum over ordered edges (c, x) of
                          //   C(d_c - 1, 2)(d_x - 1) - t_cx (d_c - 2)
  kCensusCricket = 16,    // sum over triangles, vertices of C(d - 2, 2)
  kCensusBullDeg = 17,    // sum over triangles, vertex pairs of
                          //   (d_a - 2)(d_b - 2)
  kCensusTriS = 18,       // sum over vertices of t_v s_v,
                          //   s_v = sum over neighbours x of (d_x - 1)
  kCensusDp3 = 19,        // sum over edges of C(t_e, 2)(d_a + d_b - 6)
  kCensusTv2 = 20,        // sum over vertices of C(t_v, 2)
  kCensusBanner = 21,     // sum over 4-cycles, vertices of (d_v - 2)
  kCensusCycEdgeT = 22,   // sum over edges of t_e q_e (q_e: 4-cycles on e)
  kCensusDp2 = 23,        // sum over triangles, edges of (t_e - 1)(d_opp - 2)
  kCensusGem = 24,        // sum over triangles, vertices h of
                          //   (t_hx - 1)(t_hy - 1)
  kCensusW4 = 25,         // wheels: sum over h of 4-cycles inside N(h)
  kCensusK23 = 26,        // sum over vertex pairs of C(codeg, 3)
  kCensusK23e = 27,       // sum over vertex pairs {a,c} of
                          //   E(N(a) ∩ N(c)) (codeg(a,c) - 2)
  kCensusTableMiss = 28,  // pair lookups that found no entry (0)
  kCensusSlots = 32
};

// The census runs on the graph relabelled in DAG order (census_relabel):
// rank is id, so "ranked below v" is "id below v", every CSR row is sorted
// with the vertices below its own first, and the DAG rows are the tails.
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
  // 4-cycle pass (census_c4_launch): undirected CSR (relabelled, rows
  // sorted), and per workgroup a
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
  // 5-vertex census extras (P2, P3, P5, P6): the terms above marked full5
  // are accumulated only when set; t_csr holds t_e per undirected CSR arc
  // (census_t_csr) for the 4-cycle pass, tv twice the triangles per vertex.
  bool full5 = false;
  const uint32_t *t_csr = nullptr;
  unsigned long long *tv = nullptr;
};

// The pair passes of the 5-vertex census (census5_kernels.hip).  Anchors
// are the vertices, each processed as the lowest-ranked vertex (DAG order:
// degree, then id) of the occurrences it accounts for, in classes by DAG
// out-degree d+ (kC5Classes; class k < kC5Classes - 1 holds d+ up to
// 64 * kC5Words[k] and keeps the adjacency inside N+(a) in LDS, the last
// class the rest, with it in a per-workgroup global slab).
constexpr int kC5Classes = 5;
constexpr int kC5Words[kC5Classes - 1] = {1, 2, 4, 8};

struct Census5Args {
  const int64_t *rowptr = nullptr;   // undirected CSR
  const int32_t *col = nullptr;
  const int32_t *src = nullptr;      // row of every CSR arc (nnz)
  const int32_t *deg = nullptr;      // n
  int32_t n = 0;
  int64_t nnz = 0;
  const int64_t *drp = nullptr;      // DAG
  const int32_t *dcol = nullptr;
  const uint32_t *t = nullptr;       // per DAG arc triangles
  uint32_t *t_csr = nullptr;         // out: per CSR arc triangles (nnz)
  const unsigned long long *tv = nullptr;  // twice the triangles per vertex
  // Anchor classes: ids grouped by class (n), class sizes and the starts
  // of the classes in `anchors` (kC5Classes each, zeroed before
  // classification).
  int32_t *anchors = nullptr;
  unsigned *cls_count = nullptr;
  unsigned *cls_fill = nullptr;
  // Per-workgroup dense arrays over all vertices: stamps (n u32, zeroed
  // once), pending counts lo (n u32) and masks (n x words u64) (zeroed,
  // and left zeroed), and a list of the vertices touched (n int32).
  uint32_t *stamp = nullptr;
  uint32_t *lo = nullptr;
  unsigned long long *mask = nullptr;
  int32_t *list = nullptr;
  int groups = 0;                    // workgroups stamp/lo/list hold
  int64_t mask_words = 0;            // mask buffer length / n (>= words of
                                     // every class)
  int gen_words = 1;                 // words per mask of the last class
  unsigned long long *grows = nullptr;  // last class: rows per workgroup
  int64_t grows_len = 0;             // max d+ x gen_words
  // Pair table: codeg(p, q) of every pair (p below q) with a common
  // neighbour below p, hashed per p into tab_off[p] .. tab_off[p + 1]
  // (a power of two of slots, or none).
  unsigned long long *tab_cap = nullptr;  // n + 1 (out of the count pass)
  unsigned long long *tab_off = nullptr;  // n + 1 (exclusive scan of cap)
  unsigned long long *table = nullptr;    // zeroed
  unsigned *cursor = nullptr;        // 16 counters (zeroed)
  unsigned long long *acc = nullptr; // kCensusSlots
};

// Enqueues t_csr and the per-vertex sums (kCensusTriS, kCensusTv2); needs
// t and tv from census_launch.
int census5_flat_launch(const Census5Args &c, int cus, hipStream_t s);
// The graph relabelled in DAG order: vertex v becomes its rank in the
// order by (degree, id), so that DAG order is id order -- row v of
// (out_rowptr, out_col) is sorted, its first d(v) - d+(v) entries lie
// below v and the rest is N+(v).  Every census sum is a count of subgraphs,
// the same on either labelling.  One stable sort of the vertices by degree
// gives the ranks, one sort of the arcs by (rank of the row, rank of the
// head) the rows.  out_rowptr holds n + 1, out_col nnz words; tmp holds
// census_relabel_bytes(n, nnz) bytes.  Returns 0, or -1 if a launch, scan
// or sort failed.
size_t census_relabel_bytes(int32_t n, int64_t nnz);
int census_relabel(const int64_t *rowptr, const int32_t *col, int32_t n,
                   int64_t nnz, int64_t *out_rowptr, int32_t *out_col,
                   void *tmp, size_t tmp_bytes, int cus, hipStream_t s);
// Enqueues the pair-table count pass (tab_cap, from zeroed stamps); the
// caller scans tab_cap into tab_off and sizes the table from tab_off[n].
int census5_table_count_launch(const Census5Args &c, hipStream_t s);
// Enqueues the classification of the anchors (cls_count, cls_fill zeroed).
int census5_classify_launch(const Census5Args &c, hipStream_t s);
// Exclusive prefix sum of `count` uint64 values; tmp of census5_scan_bytes.
int census5_scan_bytes(int64_t count, size_t *bytes);
int census5_scan(const unsigned long long *in, unsigned long long *out,
                 int64_t count, void *tmp, size_t tmp_bytes, hipStream_t s);
// Enqueues the anchor and out-neighbourhood passes of every class
// (kCensusW4, kCensusK23, kCensusK23e, kCensusTableMiss).
int census5_pairs_launch(const Census5Args &c, hipStream_t s);
// Per-workgroup bytes of the dense arrays with masks of `words` words.
size_t census5_group_bytes(int32_t n, int words);

// Waves of the enumeration kernel (slab rows to allocate).
int census_waves(int cus);

// Enqueues the source of every DAG arc into dsrc.  Returns 0, or -1 if
// the launch failed.
int census_dag_sources(const int64_t *drp, int32_t n, int32_t *dsrc, int cus,
                       hipStream_t s);

// Enqueues the enumeration and the sums (needs dsrc).  Returns 0, or -1 if
// a launch failed.
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

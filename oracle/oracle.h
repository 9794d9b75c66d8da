/*
 * oracle.h — CPU restatement of the reference's counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the
 * parity reference / CPU baseline; the product (mercury_amd) never calls it.
 *
 * All graphs are simple undirected CSR (sorted rows, no self-loops, no
 * duplicates, symmetric): rowptr int64[n+1], colidx int32[rowptr[n]].
 *
 * How each function is pinned (checked independently of its own formula):
 *   orc_orientation  - against the reference's own graph.cc compiled from
 *                      its sources (oracle/_ref) and DAG invariants.
 *   orc_tc, orc_tc_dag - brute-force triangle enumeration and closed forms
 *                      (K_n, wheels, K_{a,b}, Petersen, trees).
 *   orc_pattern      - brute-force induced-subgraph enumeration
 *                      (orc_brute) combined with the multiplicities the
 *                      reference's set algebra implies, plus closed forms on
 *                      stars, paths, cycles and complete graphs.
 *   orc_brute        - exhaustive subset enumeration on n <= 64; checked
 *                      against closed forms (K_n, K_{1,k}, C_k, P_k).
 */
#ifndef MGPM_ORACLE_H
#define MGPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* graph.cc:196-321 orientation_with_division(deg_thd).  Returns the DAG's
 * max out-degree; sizes[0..2] = DAG / small / large arc counts.  Array
 * arguments may be NULL (sizes only). */
int32_t orc_orientation(int32_t n, const int64_t *rowptr,
                        const int32_t *colidx, int32_t deg_thd,
                        int64_t sizes[3], int64_t *dag_rowptr,
                        int32_t *dag_colidx, int64_t *small_rowptr,
                        int32_t *small_colidx, int64_t *large_rowptr,
                        int32_t *large_colidx);

/* tc_challenge.cu:11-118 + gpu_kernels/tc_challenge.cuh:4-303: DAG and
 * small/large division (DEG_THD = 128), then the sum over small edges and
 * over large edges of |N+(v) ∩ N+(u)| by sorted merge. */
uint64_t orc_tc(int32_t n, const int64_t *rowptr, const int32_t *colidx,
                int32_t nthreads);

/* Counting phase only, on a given DAG: sum over arcs u->v of
 * |N+(u) ∩ N+(v)| by sorted merge, OpenMP dynamic over vertices.  This is
 * the CPU baseline. */
uint64_t orc_tc_dag(int32_t n, const int64_t *dag_rowptr,
                    const int32_t *dag_colidx, int32_t nthreads);

/* Same as orc_tc_dag restricted to source vertices [v_lo, v_hi) (bounded
 * CPU-baseline samples). */
uint64_t orc_tc_dag_range(int32_t n, const int64_t *dag_rowptr,
                          const int32_t *dag_colidx, int32_t v_lo,
                          int32_t v_hi, int32_t nthreads);

/* subgraph_matching/pattern_kernels.cuh:6-963 set algebra for pattern
 * 1..8 over every directed edge in CSR order (graph.cc:336-365), combined
 * as sm_multigpu.cu:204-233 does for one rank.  counters[0] / [1] receive
 * the kernels' counters[1] / counters[2]; returns the combined count. */
uint64_t orc_pattern(int32_t n, const int64_t *rowptr, const int32_t *colidx,
                     int32_t pattern, int32_t nthreads, uint64_t counters[2]);

/* Same, restricted to source vertices [v_lo, v_hi) (for P1 also the
 * frequency pass).  Used for bounded CPU-baseline samples. */
uint64_t orc_pattern_range(int32_t n, const int64_t *rowptr,
                           const int32_t *colidx, int32_t pattern,
                           int32_t v_lo, int32_t v_hi, int32_t nthreads,
                           uint64_t counters[2]);
/* orc_pattern with the set-op work W = sum over set operations of
 * (|a| + |b|) (membership tests as |a| = 1) in *work: the quantity the
 * reference's PROFILE macro instruments (tuning_schedules.cuh:32-39). */
uint64_t orc_pattern_work(int32_t n, const int64_t *rowptr,
                          const int32_t *colidx, int32_t pattern,
                          int32_t nthreads, uint64_t counters[2],
                          uint64_t *work);

/* Exhaustive induced-subgraph census on n <= 64 vertices.  out[] gets:
 *  0 triangles            1 claw K1,3 (induced)   2 path P4 (induced)
 *  3 paw (induced)        4 K1,4                  5 fork/chair
 *  6 bull                 7 diamond + pendant at a degree-3 vertex
 *  8 K1,1,3 (book)        9 K4 + pendant          10 4-cliques
 *  11 5-cliques
 * Returns 0, or -1 if n > 64. */
int32_t orc_brute(int32_t n, const int64_t *rowptr, const int32_t *colidx,
                  uint64_t out[12]);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.c — CPU restatement of the reference's counting path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the parity reference and the CPU
 * baseline, never the product.  Set algebra is done by sorted merges, so
 * it shares no machinery with the GPU kernels (hash tables, binary search).
 * All counts are exact uint64.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define DEG(v) ((int32_t)(rp[(v) + 1] - rp[(v)]))
#define NB(v) (ci + rp[(v)])

/* ------------------------------------------------------------------ sets */

/* Set-op work of the pattern passes, per thread: the sum over set
 * operations of (|a| + |b|) (a membership test is a set operation with
 * |a| = 1), the quantity the reference's PROFILE macro instruments
 * (gpu_kernels/tuning_schedules.cuh:32-39). */
static _Thread_local uint64_t g_work;

/* exact membership in a sorted list */
static int has(const int32_t *a, int32_t na, int32_t key) {
  g_work += 1u + (uint64_t)na;
  int32_t lo = 0, hi = na;
  while (lo < hi) {
    int32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < na && a[lo] == key;
}

/* {x in a : x not in b, x < ub} by merge (ub = INT32_MAX: no bound).
 * set_difference.cuh:129-163 / 200-237 semantics. */
static int32_t diff_ub(const int32_t *a, int32_t na, const int32_t *b,
                       int32_t nb, int32_t ub, int32_t *out) {
  g_work += (uint64_t)na + (uint64_t)nb;
  int32_t i = 0, j = 0, k = 0;
  while (i < na && a[i] < ub) {
    while (j < nb && b[j] < a[i]) ++j;
    if (!(j < nb && b[j] == a[i])) {
      if (out) out[k] = a[i];
      ++k;
    }
    ++i;
  }
  return k;
}

/* {x in a and b : x < ub} by merge.  set_intersect.cuh:73-193 semantics. */
static int32_t inter_ub(const int32_t *a, int32_t na, const int32_t *b,
                        int32_t nb, int32_t ub, int32_t *out) {
  g_work += (uint64_t)na + (uint64_t)nb;
  int32_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (a[i] >= ub || b[j] >= ub) break;
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      if (out) out[k] = a[i];
      ++k;
      ++i;
      ++j;
    }
  }
  return k;
}

#define NOUB INT32_MAX

/* ----------------------------------------------------------- orientation */

static int oriented(const int64_t *rp, int32_t s, int32_t t) {
  /* graph.cc:207-213 / 226-232: keep s->t iff deg(t) > deg(s) or equal
   * degrees and t > s (deg = undirected degree). */
  const int32_t ds = DEG(s), dt = DEG(t);
  return dt > ds || (dt == ds && t > s);
}

int32_t orc_orientation(int32_t n, const int64_t *rp, const int32_t *ci,
                        int32_t deg_thd, int64_t sizes[3], int64_t *dag_rowptr,
                        int32_t *dag_colidx, int64_t *small_rowptr,
                        int32_t *small_colidx, int64_t *large_rowptr,
                        int32_t *large_colidx) {
  int32_t *nd = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  int64_t *sd = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *ld = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int32_t maxd = 0;
  /* graph.cc:205-214: DAG out-degree */
  for (int32_t s = 0; s < n; ++s) {
    int32_t c = 0;
    for (int64_t e = rp[s]; e < rp[s + 1]; ++e) c += oriented(rp, s, ci[e]);
    nd[s] = c;
    if (c > maxd) maxd = c;
  }
  /* graph.cc:237-264: classify every undirected adjacency entry */
  for (int32_t s = 0; s < n; ++s) {
    for (int64_t e = rp[s]; e < rp[s + 1]; ++e) {
      const int32_t t = ci[e];
      const int32_t a = nd[s], b = nd[t];
      if ((a < b ? a : b) < 1) continue;
      if ((a > b ? a : b) <= deg_thd) {
        if (oriented(rp, s, t)) sd[s]++;
      } else if (a >= deg_thd || b >= deg_thd) {
        /* REVERSE_DAG (graph.cc:4, 253-258) */
        if (a > b || (a == b && s > t)) ld[s]++;
      }
    }
  }
  int64_t tot[3] = {0, 0, 0};
  for (int32_t s = 0; s < n; ++s) {
    if (dag_rowptr) dag_rowptr[s] = tot[0];
    if (small_rowptr) small_rowptr[s] = tot[1];
    if (large_rowptr) large_rowptr[s] = tot[2];
    tot[0] += nd[s];
    tot[1] += sd[s];
    tot[2] += ld[s];
  }
  if (dag_rowptr) dag_rowptr[n] = tot[0];
  if (small_rowptr) small_rowptr[n] = tot[1];
  if (large_rowptr) large_rowptr[n] = tot[2];
  sizes[0] = tot[0];
  sizes[1] = tot[1];
  sizes[2] = tot[2];
  /* graph.cc:221-235 and 265-312: fill in adjacency order */
  int64_t o0 = 0, o1 = 0, o2 = 0;
  for (int32_t s = 0; s < n; ++s) {
    for (int64_t e = rp[s]; e < rp[s + 1]; ++e) {
      const int32_t t = ci[e];
      if (oriented(rp, s, t)) {
        if (dag_colidx) dag_colidx[o0] = t;
        ++o0;
      }
      const int32_t a = nd[s], b = nd[t];
      if ((a < b ? a : b) < 1) continue;
      if ((a > b ? a : b) <= deg_thd) {
        if (oriented(rp, s, t)) {
          if (small_colidx) small_colidx[o1] = t;
          ++o1;
        }
      } else if (a >= deg_thd || b >= deg_thd) {
        if (a > b || (a == b && s > t)) {
          if (large_colidx) large_colidx[o2] = t;
          ++o2;
        }
      }
    }
  }
  free(nd);
  free(sd);
  free(ld);
  return maxd;
}

/* -------------------------------------------------------------------- TC */

static uint64_t merge_count(const int32_t *a, int64_t na, const int32_t *b,
                            int64_t nb) {
  uint64_t c = 0;
  int64_t i = 0, j = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      ++c;
      ++i;
      ++j;
    }
  }
  return c;
}

uint64_t orc_tc(int32_t n, const int64_t *rp, const int32_t *ci,
                int32_t nthreads) {
  int64_t sz[3];
  orc_orientation(n, rp, ci, 128, sz, NULL, NULL, NULL, NULL, NULL, NULL);
  int64_t *drp = (int64_t *)malloc(((size_t)n + 1) * 8);
  int64_t *srp = (int64_t *)malloc(((size_t)n + 1) * 8);
  int64_t *lrp = (int64_t *)malloc(((size_t)n + 1) * 8);
  int32_t *dci = (int32_t *)malloc((size_t)(sz[0] + 1) * 4);
  int32_t *sci = (int32_t *)malloc((size_t)(sz[1] + 1) * 4);
  int32_t *lci = (int32_t *)malloc((size_t)(sz[2] + 1) * 4);
  orc_orientation(n, rp, ci, 128, sz, drp, dci, srp, sci, lrp, lci);
  uint64_t total = 0;
  if (nthreads < 1) nthreads = omp_get_max_threads();
  /* binarySearch_for_small_degree: v with small edges, u in small(v);
   * hashIndex_for_large_degree: v with large edges, u in large(v);
   * both add |N+(v) ∩ N+(u)| (tc_challenge.cuh:97-183, 242-292). */
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) \
    num_threads(nthreads)
  for (int32_t v = 0; v < n; ++v) {
    const int32_t *nv = dci + drp[v];
    const int64_t dv = drp[v + 1] - drp[v];
    for (int64_t e = srp[v]; e < srp[v + 1]; ++e) {
      const int32_t u = sci[e];
      total += merge_count(nv, dv, dci + drp[u], drp[u + 1] - drp[u]);
    }
    for (int64_t e = lrp[v]; e < lrp[v + 1]; ++e) {
      const int32_t u = lci[e];
      total += merge_count(nv, dv, dci + drp[u], drp[u + 1] - drp[u]);
    }
  }
  free(drp);
  free(srp);
  free(lrp);
  free(dci);
  free(sci);
  free(lci);
  return total;
}

uint64_t orc_tc_dag_range(int32_t n, const int64_t *drp, const int32_t *dci,
                          int32_t v_lo, int32_t v_hi, int32_t nthreads) {
  uint64_t total = 0;
  if (v_lo < 0) v_lo = 0;
  if (v_hi > n) v_hi = n;
  if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) \
    num_threads(nthreads)
  for (int32_t u = v_lo; u < v_hi; ++u) {
    const int32_t *nu = dci + drp[u];
    const int64_t du = drp[u + 1] - drp[u];
    for (int64_t e = 0; e < du; ++e) {
      const int32_t v = nu[e];
      total += merge_count(nu, du, dci + drp[v], drp[v + 1] - drp[v]);
    }
  }
  return total;
}

uint64_t orc_tc_dag(int32_t n, const int64_t *drp, const int32_t *dci,
                    int32_t nthreads) {
  return orc_tc_dag_range(n, drp, dci, 0, n, nthreads);
}

/* -------------------------------------------------------------- patterns */

typedef struct {
  int32_t *l[8];
  int64_t *f;
} scratch_t;

/* P1_frequency_count (pattern_kernels.cuh:68-133) for vertex v0 -> c1. */
static uint64_t p1_frequency(const int64_t *rp, const int32_t *ci, int32_t v0,
                             scratch_t *s) {
  const int32_t d = DEG(v0);
  const int32_t *n0 = NB(v0);
  int64_t *f = s->f;
  for (int32_t i = 0; i < d; ++i) f[i] = 0;
  for (int32_t j = 0; j < d; ++j) {
    const int32_t v1 = n0[j];
    for (int32_t i = 0; i < d; ++i) {
      const int32_t key = n0[i];
      if (key < v1 && !has(NB(v1), DEG(v1), key)) f[i]++;
    }
  }
  uint64_t c1 = 0;
  for (int32_t k = 0; k < d; ++k) {
    const int32_t v2 = n0[k];
    const uint64_t cnt = (uint64_t)diff_ub(n0, d, NB(v2), DEG(v2), v2, NULL);
    c1 += cnt * (uint64_t)f[k];
  }
  return c1;
}

/* Pattern kernels for one directed edge (v0, v1); adds to c[0] (the
 * kernel's counters[1]) and c[1] (counters[2]). */
static void pattern_edge(const int64_t *rp, const int32_t *ci, int pat,
                         int32_t v0, int32_t v1, scratch_t *s, uint64_t c[2]) {
  const int32_t *n0 = NB(v0), *n1 = NB(v1);
  const int32_t d0 = DEG(v0), d1 = DEG(v1);
  int32_t **L = s->l;
  int64_t *f = s->f;
  switch (pat) {
    case 1: { /* P1_count_correction :6-66 */
      if (v0 == v1) return;
      const int32_t na = diff_ub(n0, d0, n1, d1, v1, L[0]);
      const int32_t nb = inter_ub(n0, d0, n1, d1, v1, L[1]);
      for (int32_t i = 0; i < nb; ++i) {
        const int32_t v2 = L[1][i];
        for (int32_t j = 0; j < na; ++j) {
          const int32_t key = L[0][j];
          if (key > v2 && !has(NB(key), DEG(key), v2)) c[1]++;
        }
      }
      return;
    }
    case 2: { /* P2_subgraph_matching :135-243 */
      const int32_t n_l0 = diff_ub(n0, d0, n1, d1, NOUB, L[0]);
      const int32_t n_l1 = diff_ub(n1, d1, n0, d0, NOUB, L[1]);
      for (int32_t j = 0; j < n_l0; ++j) f[j] = 0;
      for (int32_t i = 0; i < n_l1; ++i) {
        const int32_t v4 = L[1][i];
        for (int32_t j = 0; j < n_l0; ++j)
          if (!has(NB(v4), DEG(v4), L[0][j])) f[j]++;
      }
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        c[0] += (uint64_t)diff_ub(L[0], n_l0, NB(v2), DEG(v2), v2, NULL) *
                (uint64_t)f[i];
      }
      for (int32_t i = 0; i < n_l1; ++i) {
        const int32_t v4 = L[1][i];
        const int32_t n2 = diff_ub(L[0], n_l0, NB(v4), DEG(v4), NOUB, L[2]);
        const int32_t n3 = inter_ub(L[0], n_l0, NB(v4), DEG(v4), NOUB, L[3]);
        for (int32_t ii = 0; ii < n3; ++ii) {
          const int32_t v2 = L[3][ii];
          for (int32_t j = 0; j < n2; ++j) {
            const int32_t key = L[2][j];
            if (key > v2 && !has(NB(key), DEG(key), v2)) c[1]++;
          }
        }
      }
      return;
    }
    case 3:   /* P3_subgraph_matching :245-350 (L0 = N0 \ N1) */
    case 7: { /* P7_subgraph_matching :352-459 (L0 = N0 ∩ N1) */
      const int32_t n_l0 = pat == 3 ? diff_ub(n0, d0, n1, d1, NOUB, L[0])
                                    : inter_ub(n0, d0, n1, d1, NOUB, L[0]);
      for (int32_t j = 0; j < n_l0; ++j) f[j] = 0;
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        for (int32_t j = 0; j < i; ++j)
          if (!has(NB(v2), DEG(v2), L[0][j])) f[j]++;
      }
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        c[0] += (uint64_t)diff_ub(L[0], n_l0, NB(v2), DEG(v2), v2, NULL) *
                (uint64_t)f[i];
      }
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t y = L[0][i];
        const int32_t n1l = diff_ub(L[0], n_l0, NB(y), DEG(y), y, L[1]);
        const int32_t n2l = inter_ub(L[0], n_l0, NB(y), DEG(y), NOUB, L[2]);
        for (int32_t ii = 0; ii < n2l; ++ii) {
          const int32_t v2 = L[2][ii];
          for (int32_t j = 0; j < n1l; ++j) {
            const int32_t key = L[1][j];
            if (key > v2 && !has(NB(key), DEG(key), v2)) c[1]++;
          }
        }
      }
      return;
    }
    case 8: { /* P8_subgraph_matching :461-572 */
      const int32_t n_l0 = inter_ub(n0, d0, n1, d1, NOUB, L[0]);
      const int32_t n_l1 = diff_ub(n1, d1, n0, d0, NOUB, L[1]);
      for (int32_t j = 0; j < n_l0; ++j) f[j] = 0;
      for (int32_t i = 0; i < n_l1; ++i) {
        const int32_t v4 = L[1][i];
        for (int32_t j = 0; j < n_l0; ++j)
          if (!has(NB(v4), DEG(v4), L[0][j])) f[j]++;
      }
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        c[0] += (uint64_t)inter_ub(L[0], n_l0, NB(v2), DEG(v2), v2, NULL) *
                (uint64_t)f[i];
      }
      for (int32_t i = 0; i < n_l1; ++i) {
        const int32_t v4 = L[1][i];
        const int32_t n2 = diff_ub(L[0], n_l0, NB(v4), DEG(v4), NOUB, L[2]);
        const int32_t n3 = inter_ub(L[0], n_l0, NB(v4), DEG(v4), NOUB, L[3]);
        for (int32_t ii = 0; ii < n3; ++ii) {
          const int32_t v2 = L[3][ii];
          for (int32_t j = 0; j < n2; ++j) {
            const int32_t key = L[2][j];
            if (key > v2 && has(NB(key), DEG(key), v2)) c[1]++;
          }
        }
      }
      return;
    }
    case 5: { /* P5_subgraph_matching :574-734 */
      const int32_t n_l0 = inter_ub(n0, d0, n1, d1, NOUB, L[0]);
      const int32_t n_l1 = diff_ub(n0, d0, n1, d1, NOUB, L[1]);
      const int32_t n_l2 = diff_ub(n1, d1, n0, d0, NOUB, L[2]);
      for (int32_t j = 0; j < n_l1; ++j) f[j] = 0;
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        for (int32_t j = 0; j < n_l1; ++j)
          if (!has(NB(v2), DEG(v2), L[1][j])) f[j]++;
      }
      for (int32_t i = 0; i < n_l1; ++i) {
        if (f[i] == 0) continue;
        const int32_t v2 = L[1][i];
        uint64_t t = 0;
        for (int32_t j = 0; j < n_l2; ++j)
          if (!has(NB(v2), DEG(v2), L[2][j])) ++t;
        c[0] += t * (uint64_t)f[i];
      }
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t v2 = L[0][i];
        const int32_t n5 = inter_ub(L[2], n_l2, NB(v2), DEG(v2), NOUB, L[5]);
        const int32_t n4 = diff_ub(L[1], n_l1, NB(v2), DEG(v2), NOUB, L[4]);
        for (int32_t j = 0; j < n5; ++j) {
          const int32_t b = L[5][j];
          for (int32_t k = 0; k < n4; ++k)
            if (!has(NB(b), DEG(b), L[4][k])) c[1]++;
        }
      }
      return;
    }
    case 6: { /* P6_subgraph_matching :736-853 */
      const int32_t n_l0 = inter_ub(n0, d0, n1, d1, NOUB, L[0]);
      const int32_t n_l1 = diff_ub(n1, d1, n0, d0, NOUB, L[1]);
      for (int32_t i = 0; i < n_l0; ++i) {
        const int32_t cc = L[0][i];
        const int32_t n2 = diff_ub(L[0], n_l0, NB(cc), DEG(cc), NOUB, L[2]);
        const int32_t n3 = diff_ub(L[1], n_l1, NB(cc), DEG(cc), NOUB, L[3]);
        /* both branches of :795-850 count pairs (d,e) in L2 x L3, d !~ e */
        for (int32_t j = 0; j < n2; ++j) {
          const int32_t d = L[2][j];
          for (int32_t k = 0; k < n3; ++k)
            if (!has(NB(d), DEG(d), L[3][k])) c[0]++;
        }
      }
      return;
    }
    case 4: { /* P4_subgraph_matching :855-963, only v1 < v0 */
      if (v1 >= v0) return;
      const int32_t n_l0 = diff_ub(n0, d0, n1, d1, NOUB, L[0]);
      const int32_t n_l1 = diff_ub(n1, d1, n0, d0, NOUB, L[1]);
      for (int32_t j = 0; j < n_l0; ++j) {
        const int32_t a = L[0][j];
        for (int32_t k = 0; k < n_l1; ++k)
          if (!has(NB(a), DEG(a), L[1][k])) c[0]++;
      }
      return;
    }
    default:
      return;
  }
}

static uint64_t pattern_range(int32_t n, const int64_t *rp, const int32_t *ci,
                              int32_t pat, int32_t v_lo, int32_t v_hi,
                              int32_t nthreads, uint64_t counters[2],
                              uint64_t *work);

uint64_t orc_pattern_range(int32_t n, const int64_t *rp, const int32_t *ci,
                           int32_t pat, int32_t v_lo, int32_t v_hi,
                           int32_t nthreads, uint64_t counters[2]) {
  return pattern_range(n, rp, ci, pat, v_lo, v_hi, nthreads, counters, NULL);
}

uint64_t orc_pattern_work(int32_t n, const int64_t *rp, const int32_t *ci,
                          int32_t pat, int32_t nthreads, uint64_t counters[2],
                          uint64_t *work) {
  return pattern_range(n, rp, ci, pat, 0, n, nthreads, counters, work);
}

static uint64_t pattern_range(int32_t n, const int64_t *rp, const int32_t *ci,
                              int32_t pat, int32_t v_lo, int32_t v_hi,
                              int32_t nthreads, uint64_t counters[2],
                              uint64_t *work) {
  if (pat < 1 || pat > 8) return 0;
  if (v_lo < 0) v_lo = 0;
  if (v_hi > n) v_hi = n;
  int32_t maxd = 1;
  for (int32_t v = 0; v < n; ++v)
    if (DEG(v) > maxd) maxd = DEG(v);
  if (nthreads < 1) nthreads = omp_get_max_threads();
  uint64_t c1 = 0, c2 = 0, w = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : c1, c2, w)
  {
    g_work = 0;
    scratch_t s;
    for (int k = 0; k < 8; ++k)
      s.l[k] = (int32_t *)malloc((size_t)maxd * sizeof(int32_t));
    s.f = (int64_t *)malloc((size_t)maxd * sizeof(int64_t));
#pragma omp for schedule(dynamic, 16)
    for (int32_t v0 = v_lo; v0 < v_hi; ++v0) {
      uint64_t c[2] = {0, 0};
      /* sm_multigpu.cu:150-155: P1 runs the frequency pass over vertices
       * and the correction over directed edges. */
      if (pat == 1) c[0] += p1_frequency(rp, ci, v0, &s);
      for (int64_t e = rp[v0]; e < rp[v0 + 1]; ++e)
        pattern_edge(rp, ci, pat, v0, ci[e], &s, c);
      c1 += c[0];
      c2 += c[1];
    }
    for (int k = 0; k < 8; ++k) free(s.l[k]);
    free(s.f);
    w += g_work;
  }
  if (counters) {
    counters[0] = c1;
    counters[1] = c2;
  }
  if (work) *work = w;
  /* sm_multigpu.cu:209-232 (one rank): P4, P6 -> c1; others c1 - c2 */
  if (pat == 4 || pat == 6) return c1;
  return c1 - c2;
}

uint64_t orc_pattern(int32_t n, const int64_t *rp, const int32_t *ci,
                     int32_t pat, int32_t nthreads, uint64_t counters[2]) {
  return orc_pattern_range(n, rp, ci, pat, 0, n, nthreads, counters);
}

/* ---------------------------------------------------------- brute force */

static int cmp_desc(const void *a, const void *b) {
  return *(const int *)b - *(const int *)a;
}

static int same(const int *d, const int *e, int k) {
  for (int i = 0; i < k; ++i)
    if (d[i] != e[i]) return 0;
  return 1;
}

int32_t orc_brute(int32_t n, const int64_t *rp, const int32_t *ci,
                  uint64_t out[12]) {
  if (n > 64) return -1;
  uint64_t adj[64];
  memset(adj, 0, sizeof(adj));
  for (int32_t v = 0; v < n; ++v)
    for (int64_t e = rp[v]; e < rp[v + 1]; ++e) adj[v] |= 1ull << ci[e];
  for (int i = 0; i < 12; ++i) out[i] = 0;
  static const int claw[4] = {3, 1, 1, 1}, path4[4] = {2, 2, 1, 1},
                   paw[4] = {3, 2, 2, 1}, k4[4] = {3, 3, 3, 3};
  static const int k14[5] = {4, 1, 1, 1, 1}, fork5[5] = {3, 2, 1, 1, 1},
                   bull[5] = {3, 3, 2, 1, 1}, dpend[5] = {4, 3, 2, 2, 1},
                   k113[5] = {4, 4, 2, 2, 2}, k4p[5] = {4, 3, 3, 3, 1},
                   k5[5] = {4, 4, 4, 4, 4};
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b)
      for (int c = b + 1; c < n; ++c) {
        const uint64_t s3 = (1ull << a) | (1ull << b) | (1ull << c);
        if (__builtin_popcountll(adj[a] & s3) == 2 &&
            __builtin_popcountll(adj[b] & s3) == 2)
          out[0]++;
        for (int d = c + 1; d < n; ++d) {
          const uint64_t s4 = s3 | (1ull << d);
          int dg4[4] = {__builtin_popcountll(adj[a] & s4),
                        __builtin_popcountll(adj[b] & s4),
                        __builtin_popcountll(adj[c] & s4),
                        __builtin_popcountll(adj[d] & s4)};
          qsort(dg4, 4, sizeof(int), cmp_desc);
          if (same(dg4, claw, 4)) out[1]++;
          else if (same(dg4, path4, 4)) out[2]++;
          else if (same(dg4, paw, 4)) out[3]++;
          else if (same(dg4, k4, 4)) out[10]++;
          for (int e = d + 1; e < n; ++e) {
            const uint64_t s5 = s4 | (1ull << e);
            int dg[5] = {__builtin_popcountll(adj[a] & s5),
                         __builtin_popcountll(adj[b] & s5),
                         __builtin_popcountll(adj[c] & s5),
                         __builtin_popcountll(adj[d] & s5),
                         __builtin_popcountll(adj[e] & s5)};
            qsort(dg, 5, sizeof(int), cmp_desc);
            if (same(dg, k14, 5)) out[4]++;
            else if (same(dg, fork5, 5)) out[5]++;
            else if (same(dg, bull, 5)) out[6]++;
            else if (same(dg, dpend, 5)) out[7]++;
            else if (same(dg, k113, 5)) out[8]++;
            else if (same(dg, k4p, 5)) out[9]++;
            else if (same(dg, k5, 5)) out[11]++;
          }
        }
      }
  return 0;
}

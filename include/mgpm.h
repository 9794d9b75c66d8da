/*
 * mgpm.h — C ABI of the MI355X-native MERCURY counting path.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns
 * MGPM_OK (0) or a negative MGPM_ERR_* code; nothing here calls exit() or
 * throws across the boundary.  The caller owns every handle it receives and
 * frees it with the matching *_free / *_destroy call.  One handle is used by
 * one host thread at a time.  Counts are uint64 (exact, wrapping mod 2^64
 * exactly where the reference's unsigned long long accumulators wrap).
 *
 * Streams are passed as `void*` holding a hipStream_t (NULL = the default
 * stream of the device the handle lives on).
 *
 * Reference interfaces replaced (reference file:line):
 *   main.cc:5-20                         -> the `tc_challenge` CLI over mgpm_tc_count
 *   tc_challenge.cu:11  TCSolver         -> mgpm_dgraph_upload + mgpm_tc_count
 *   subgraph_matching/main_mpi.cpp:5-38  -> the `sm` CLI over mgpm_pattern_count
 *   subgraph_matching/sm_multigpu.cu:97  SMSolverMultiGPU -> mgpm_pattern_count
 *   graph.cc:6-92  Graph::Graph + custom_alloc.h:33-44 read_file -> mgpm_graph_load
 *   graph.cc:196-321 orientation_with_division -> mgpm_graph_orientation
 *   graph_gpu.h:87-188 GraphGPU::init / init_with_division -> mgpm_dgraph_upload
 *   cutil_subset.h:4-12 CUDA_SAFE_CALL (exit on error) -> negative status codes
 */
#ifndef MGPM_H
#define MGPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPM_ABI_VERSION 1

/* ---- status codes --------------------------------------------------- */
#define MGPM_OK 0
#define MGPM_ERR_IO (-1)          /* a file could not be opened / was short */
#define MGPM_ERR_FORMAT (-2)      /* graph.meta.txt malformed or sizes wrong */
#define MGPM_ERR_INVALID (-3)     /* bad argument */
#define MGPM_ERR_NOMEM (-4)       /* host or device allocation failed */
#define MGPM_ERR_HIP (-5)         /* a HIP runtime call failed */
#define MGPM_ERR_NODEVICE (-6)    /* no usable GPU / bad device index */
#define MGPM_ERR_GRAPH (-7)       /* CSR is not simple/sorted/symmetric */
#define MGPM_ERR_PATTERN (-8)     /* unknown pattern id ("Not support.") */
#define MGPM_ERR_CAPACITY (-9)    /* request exceeds a capacity/budget */
#define MGPM_ERR_BUSY (-10)       /* resource in use */
#define MGPM_ERR_NOTFOUND (-11)   /* unknown id / ticket */
#define MGPM_ERR_TIMEOUT (-12)    /* wait timed out */
#define MGPM_ERR_FAILED (-13)     /* request failed (server) */
#define MGPM_ERR_INTERNAL (-14)

/* ---- patterns -------------------------------------------------------- */
/* MGPM_TC is triangle counting; 1..8 are the reference's P1..P8
 * (subgraph_matching/sm_multigpu.cu:148-198). */
#define MGPM_TC 0
#define MGPM_P1 1
#define MGPM_P2 2
#define MGPM_P3 3
#define MGPM_P4 4
#define MGPM_P5 5
#define MGPM_P6 6
#define MGPM_P7 7
#define MGPM_P8 8
#define MGPM_NUM_PATTERNS 9

typedef struct mgpm_graph mgpm_graph;   /* host CSR, undirected */
typedef struct mgpm_dgraph mgpm_dgraph; /* device-resident graph */
typedef struct mgpm_batch mgpm_batch;   /* batched multi-graph launcher */
typedef struct mgpm_server mgpm_server; /* continuous-batching server */

/* Knobs of the counting kernels.  Pass NULL for the defaults.  In every
 * field 0 means "default"; for the *_max_deg fields -1 empties that class.
 * Triangle tasks are classed by the table side's DAG out-degree d+:
 *   d+ <= small_max_deg  -> 256-slot wave-private LDS tables   (cap 128)
 *   d+ <= wave_max_deg   -> 2048-slot wave-private LDS tables  (cap 1024)
 *   d+ <= block_max_deg  -> workgroup-shared LDS tables        (cap 16384)
 *   larger               -> binary search in the global list. */
typedef struct mgpm_tune {
  int32_t small_max_deg;
  int32_t wave_max_deg;
  int32_t block_max_deg;
  int32_t chunk_work;     /* TC: max cost of one task (list items plus 16
                             per partner); heavier owners are split
                             (0 = 65536, doubled per doubling of the DAG's
                             arcs from 2^27 on, at most 2^20) */
  int32_t grid_blocks;    /* cap on persistent workgroups per kernel */
  int32_t pattern_waves;  /* P1..P8: waves per workgroup */
  int32_t group_work;     /* TC: light tasks are packed into queue entries
                             of up to this cost */
  int32_t tc_items;       /* TC: list items per lane per step (4 or 8;
                             default 8) */
  int32_t tc_cursor_min;  /* TC: 64-partner batches whose lists average at
                             least this many items are walked with one
                             owner cursor per lane instead of start marks
                             (default -1 = never) */
  int32_t pattern_split;  /* P1..P8: a task (edge, or P1 vertex) whose
                             estimated pair tests exceed 2^pattern_split is
                             cut into shares run by different waves
                             (default 20; -1 = never) */
  int32_t pattern_census; /* P1..P8: 0 (default) or 1 = closed form from
                             a subgraph census (triangles, 4-cycles, 4- and
                             5-cliques; for P2, P3, P5, P6 also wheels,
                             K2,3 and K2,3 plus an edge); -1 = the
                             reference's per-edge set algebra (same count) */
  int32_t tc_layout;      /* TC (mgpm_tc_count): 1 = stream every task's
                             partner lists from a per-graph copy laid out
                             task by task (built on first use and kept with
                             the plan: 4 bytes per probe + 8 per partner),
                             -1 = gather them from the DAG; 0 (default) =
                             lay out when the copy fits a quarter of the
                             free device memory.  Same count either way. */
  int32_t tc_merge_max;   /* TC: owners with d+ up to this (at most 64)
                             skip the hash table: each lane merges one
                             partner list with the owner's sorted list (two
                             cursors); 0 (default) = never, -1 = never */
  int32_t tc_half_waves;  /* TC with laid-out lists: 1 = small-class tasks
                             (d+ <= 128) two per wave, 32 lanes each with
                             its own table; -1 = one per wave; 0 (default)
                             = two per wave when every task is small */
} mgpm_tune;

/* Per-call measurements. */
typedef struct mgpm_stats {
  double kernel_ms;   /* device time of the counting kernels: hipEvents
                         recorded on the launch stream around them */
  double e2e_ms;      /* host wall time of the whole call */
  double prep_ms;     /* host/device preprocessing done inside this call
                         (0 when the plan was already resident) */
  uint64_t n;         /* vertices */
  uint64_t m;         /* undirected edges */
  uint64_t m_plus;    /* DAG arcs (TC) */
  uint64_t s_plus;    /* sum over DAG arcs u->v of d+(v) (TC wedge checks) */
  uint64_t probes;    /* TC: hash probes issued by the chosen schedule */
  uint64_t bytes_alg; /* TC: 8(n+1) + 20 m+ + 4 S+ */
  int32_t path;       /* implementation path id (diagnostic) */
  int32_t reserved[3];
} mgpm_stats;

typedef struct mgpm_graph_info {
  int32_t n;
  int32_t max_degree;      /* recomputed from the CSR */
  int64_t nnz;             /* CSR entries = 2m */
  int32_t meta_max_degree; /* as stored in graph.meta.txt (0 if not loaded) */
  int32_t reserved;
} mgpm_graph_info;

/* ---- errors / version -------------------------------------------------- */
const char *mgpm_strerror(int code);
int mgpm_abi_version(void);
/* Last error detail message of the calling thread ("" if none). */
const char *mgpm_last_error(void);

/* ---- host graphs ---------------------------------------------------------
 * Reference: graph.cc:6-92 (Graph::Graph), custom_alloc.h:33-44 (read_file).
 * Reads <prefix>.meta.txt / .vertex.bin / .edge.bin.  Where the reference
 * asserts or exits, this returns MGPM_ERR_FORMAT / MGPM_ERR_IO, and it also
 * rejects a CSR that is not simple, sorted and symmetric (MGPM_ERR_GRAPH),
 * which the reference silently assumes. */
int mgpm_graph_load(const char *prefix, mgpm_graph **out);
/* Copies the arrays.  Validated like mgpm_graph_load. */
int mgpm_graph_from_csr(int32_t n, int64_t nnz, const int64_t *rowptr,
                        const int32_t *colidx, mgpm_graph **out);
/* Edge files, in place of the reference's closed datasets/converter
 * (converter.sh:5, "converter mtx <in> <out>/graph 1 0 0 0") and
 * datasets/tsv2mtx.py.  format: 0 = by extension (".mtx" -> 1, else 2),
 * 1 = MatrixMarket coordinate, 2 = edge list, one "src dst [...]" per line
 * (GraphChallenge TSV).  base: id of the first vertex (-1 = 1, as in
 * MatrixMarket and GraphChallenge files; 0 for 0-based lists, the
 * reference's converter_vidfrom0).  Symmetrized; self-loops and duplicates
 * dropped.  Errors: MGPM_ERR_IO, MGPM_ERR_FORMAT. */
int mgpm_graph_read_edges(const char *path, int32_t format, int32_t base,
                          mgpm_graph **out);
/* Writes the reference format (meta: "n 2m 4 8 1 2 maxdeg 0 nvc nec", the
 * class counts 0 when the graph has no labels of that kind) and, when the
 * graph has labels, <prefix>.vlabel.bin / <prefix>.elabel.bin. */
int mgpm_graph_write(const mgpm_graph *g, const char *prefix);

/* ---- labels ---------------------------------------------------------------
 * Reference: graph.cc:37-82 (use_vlabel / use_elabel in Graph::Graph).
 * mgpm_graph_load plus labels: MGPM_LOAD_VLABELS reads <prefix>.vlabel.bin
 * (n uint8), MGPM_LOAD_ELABELS reads <prefix>.elabel.bin (2m uint16, one
 * per CSR entry).  As the reference asserts, it fails (MGPM_ERR_FORMAT)
 * unless 0 < num_vertex_classes < 255 and the vertex label file holds
 * exactly that many distinct labels, or unless 0 < num_edge_classes <
 * 65535 and the edge label file holds at most that many.  A missing label
 * file gets labels in 1..classes, as the reference's fallback does, drawn
 * from a hash of the vertex / entry index (deterministic; the reference
 * uses rand()).  Short label files fail with MGPM_ERR_IO. */
#define MGPM_LOAD_VLABELS 1
#define MGPM_LOAD_ELABELS 2
int mgpm_graph_load_ex(const char *prefix, int32_t flags, mgpm_graph **out);
/* Borrowed views of the labels (NULL, class count 0, when the graph has
 * none of that kind), valid until mgpm_graph_free or the next
 * mgpm_graph_set_labels.  Any output pointer may be NULL. */
int mgpm_graph_labels(const mgpm_graph *g, const uint8_t **vlabels,
                      int32_t *num_vertex_classes, const uint16_t **elabels,
                      int32_t *num_edge_classes);
/* Copies labels into g (vlabels: n entries; elabels: nnz entries; a NULL
 * array with class count 0 removes that kind), checked by the rules that
 * mgpm_graph_load_ex applies to label files.  Not concurrently with other
 * calls on g. */
int mgpm_graph_set_labels(mgpm_graph *g, int32_t num_vertex_classes,
                          const uint8_t *vlabels, int32_t num_edge_classes,
                          const uint16_t *elabels);
/* Vertices per label 0..num_vertex_classes into freq (num_vertex_classes+1
 * entries); Graph::computeLabelsFrequency (graph.cc:440-457).  Fails with
 * MGPM_ERR_FORMAT where the reference asserts (a label above
 * num_vertex_classes), MGPM_ERR_INVALID without vertex labels. */
int mgpm_graph_label_frequency(const mgpm_graph *g, int64_t *freq);
/* Neighbourhood label frequency of every vertex (Graph::BuildNLF,
 * graph.cc:483-495) as a CSR: offsets (n+1 int64), then per vertex its
 * distinct neighbour labels ascending (labels) with the number of
 * neighbours carrying each (counts).  offsets is always filled; call with
 * labels = counts = NULL first to learn offsets[n], the entries labels and
 * counts must hold.  MGPM_ERR_INVALID without vertex labels. */
int mgpm_graph_nlf(const mgpm_graph *g, int64_t *offsets, int32_t *labels,
                   int32_t *counts);
int mgpm_graph_info_get(const mgpm_graph *g, mgpm_graph_info *info);
/* Borrowed views of the CSR, valid until mgpm_graph_free. */
int mgpm_graph_csr(const mgpm_graph *g, const int64_t **rowptr,
                   const int32_t **colidx);
int mgpm_graph_free(mgpm_graph *g);

/* Synthetic graphs (no counterpart in the reference, which ships none).
 * RMAT: 2^scale vertices, edge_factor * 2^scale draws, quadrant
 * probabilities a, b, c (d = 1-a-b-c), no noise or permutation; ER:
 * `draws` uniform pairs.  Both symmetrize, drop self-loops and duplicates.
 * Draw i depends only on (seed, i), so output is independent of threads. */
int mgpm_graph_generate_rmat(int32_t scale, int32_t edge_factor, double a,
                             double b, double c, uint64_t seed,
                             mgpm_graph **out);
int mgpm_graph_generate_er(int32_t n, int64_t draws, uint64_t seed,
                           mgpm_graph **out);

/* ---- degree-oriented DAG (host) -------------------------------------------
 * Reference: graph.cc:196-321 orientation_with_division(deg_thd).  sizes[0]
 * = DAG arcs, sizes[1] = small arcs, sizes[2] = large arcs; call once with
 * NULL arrays to get the sizes, then with caller-allocated arrays (row
 * pointers n+1 int64, column arrays of the reported sizes). */
int mgpm_graph_orientation(const mgpm_graph *g, int32_t deg_thd,
                           int64_t sizes[3], int64_t *dag_rowptr,
                           int32_t *dag_colidx, int64_t *small_rowptr,
                           int32_t *small_colidx, int64_t *large_rowptr,
                           int32_t *large_colidx);

/* The same arrays built on the device from a device graph's CSR (flat
 * passes over the arcs classify them, a device-wide scan gives the row
 * offsets, an ordered compaction writes the rows), bit-equal to
 * mgpm_graph_orientation.  The device arrays stay with the handle; any
 * non-NULL host array receives a copy (row pointers n+1 int64, column
 * arrays of the reported sizes).  stats->kernel_ms is the device time of
 * the build (hipEvents on `stream`), without the copies. */
int mgpm_dgraph_orientation(mgpm_dgraph *dg, int32_t deg_thd, void *stream,
                            int64_t sizes[3], int64_t *dag_rowptr,
                            int32_t *dag_colidx, int64_t *small_rowptr,
                            int32_t *small_colidx, int64_t *large_rowptr,
                            int32_t *large_colidx, mgpm_stats *stats);

/* ---- devices and device graphs -------------------------------------------
 * Reference: graph_gpu.h:87-188.  Uploads the undirected CSR (P1..P8) and
 * builds + uploads the triangle-count plan (DAG and table/stream tasks). */
int mgpm_device_count(int *count);
int mgpm_device_synchronize(int device);
int mgpm_dgraph_upload(const mgpm_graph *g, int device, void *stream,
                       mgpm_dgraph **out);
/* mgpm_dgraph_upload with options.  MGPM_UPLOAD_DEVICE_PLAN: build the DAG
 * and the triangle-count plan on the device from the uploaded CSR (the
 * host's orientation_with_division and task lists, graph.cc:196-321 and
 * tc_challenge.cu:20-53, run on the GPU; no host preprocessing).  Counts
 * are identical; the plan's task cut and grouping follow the host rules in
 * a parallel form.  Later re-plans (tunes, autotune) stay on the device. */
#define MGPM_UPLOAD_DEVICE_PLAN 1
int mgpm_dgraph_upload_ex(const mgpm_graph *g, int device, void *stream,
                          int32_t flags, mgpm_dgraph **out);
/* Device bytes held by the handle (admission control uses this). */
int mgpm_dgraph_footprint(const mgpm_dgraph *dg, uint64_t *bytes);
int mgpm_dgraph_free(mgpm_dgraph *dg);

/* ---- counting ----------------------------------------------------------
 * Reference: tc_challenge.cu:11-118 (TCSolver) and the kernels of
 * gpu_kernels/tc_challenge.cuh:4-303. */
int mgpm_tc_count(mgpm_dgraph *dg, const mgpm_tune *tune, void *stream,
                  uint64_t *count, mgpm_stats *stats);
/* Reference: subgraph_matching/sm_multigpu.cu:97-239 and the kernels of
 * subgraph_matching/pattern_kernels.cuh:6-963, counter combination
 * sm_multigpu.cu:204-233.  pattern in 1..8 (MGPM_TC also accepted). */
int mgpm_pattern_count(mgpm_dgraph *dg, int pattern, const mgpm_tune *tune,
                       void *stream, uint64_t *count, mgpm_stats *stats);

/* ---- allocation-free count --------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the count of mgpm_tc_count (MGPM_TC)
 * or of mgpm_pattern_count's per-edge kernels (P1..P8; the same value the
 * census gives) with every scratch buffer -- queue heads, counters and the
 * per-wave frontier slab -- in one fixed library array of
 * MGPM_NOALLOC_SCRATCH_BYTES per GPU, held from first use until the
 * process ends.  No hipMalloc, hipFree or allocating library call; the
 * pattern kernels run on as many workgroups as the array holds slabs for.
 * Needs the pattern's plan resident (one earlier mgpm_pattern_count of
 * that pattern with pattern_census = -1 builds it): MGPM_ERR_INVALID
 * otherwise.  MGPM_ERR_CAPACITY if one workgroup's slab (4 waves x 8 x
 * max degree x 4 bytes) exceeds the array, i.e. max degree above 524 286.
 * Synchronises `stream`.  Calls on one GPU share its array: calls that
 * overlap on one GPU race. */
#define MGPM_NOALLOC_SCRATCH_BYTES (64u << 20)
int mgpm_pattern_count_noalloc(mgpm_dgraph *dg, int pattern, void *stream,
                               uint64_t *count);

/* ---- autotuner --------------------------------------------------------------
 * No counterpart in the reference, whose schedule constants are fixed
 * (tuning_schedules.cuh:41-47, tc_challenge.cu:67-73).  MGPM_TC:
 * coordinate descent over tc_items, tc_cursor_min, small/wave/block_max_deg
 * (hash-table class vs binary search per owner size), chunk_work and
 * group_work (task split and queue-entry packing); P1..P8: pattern_split
 * and grid_blocks.  Each candidate is timed on
 * the device (median kernel time of `reps` runs, 0 = 3) and must return the
 * same count as the starting point (MGPM_ERR_INTERNAL otherwise).  The
 * winner stays resident on the graph -- later calls with a NULL tune,
 * batches and servers use it -- and is returned in *best (may be NULL)
 * with its time in *best_ms (may be NULL). */
int mgpm_autotune(mgpm_dgraph *dg, int pattern, int32_t reps, mgpm_tune *best,
                  double *best_ms);

/* ---- batches ------------------------------------------------------------
 * No counterpart in the reference (one graph, one pattern per process).
 * Many (graph, pattern) queries run in one call with no host round trip per
 * query: one fill kernel zeroes every query's queues and counters, the
 * work is spread over `lanes` streams (longest first, each to the
 * least-loaded lane) where the counting kernels -- the same kernels as
 * mgpm_tc_count / mgpm_pattern_count, with each launch capped at the
 * workgroups its queue can feed -- overlap on the device, and one copy
 * returns all counters.  The P1..P8 queries of one graph share one census
 * run (the union of the passes they need), each combining its own count
 * from the census accumulators.  With use_graph the fill, fork, kernels and join
 * are recorded once into a HIP graph and replayed while the query set
 * (graph handles, patterns, plan generations) is unchanged.  Every query
 * uses its graph's resident plan (mgpm_autotune's choice if it ran).  All
 * graphs must live on the batch's device. */
typedef struct mgpm_query {
  mgpm_dgraph *graph; /* in */
  int32_t pattern;    /* in: MGPM_TC or MGPM_P1..MGPM_P8 */
  int32_t status;     /* out: MGPM_OK or a negative code */
  uint64_t count;     /* out */
} mgpm_query;

typedef struct mgpm_batch_opts {
  int32_t lanes;      /* concurrent streams (0 = 4, at most 16) */
  int32_t use_graph;  /* 1 = record and replay a HIP graph */
  int32_t autotune;   /* 1 = before the first run of a query set, tune the
                         triangle knobs of the largest untuned graph of each
                         size class (log2 of DAG arcs) with mgpm_autotune's
                         search and apply them to the rest of the class */
  int32_t pattern_census; /* P1..P8: 0 = closed form from the census (each
                             lane keeps a census workspace), -1 = the
                             per-edge kernels; the same counts */
  int32_t reserved[4];
} mgpm_batch_opts;

/* opts may be NULL (defaults). */
int mgpm_batch_create(int device, const mgpm_batch_opts *opts,
                      mgpm_batch **out);
/* Runs the queries; per-query results in queries[i].count / .status.
 * stats->kernel_ms is the device time of the whole batch (hipEvents on
 * `stream` around fill, kernels, join and copy). */
int mgpm_batch_run(mgpm_batch *b, mgpm_query *queries, int nq, void *stream,
                   mgpm_stats *stats);
/* Device bytes of the batch's own workspace (counters, queues, scratch). */
int mgpm_batch_footprint(const mgpm_batch *b, uint64_t *bytes);
int mgpm_batch_free(mgpm_batch *b);

/* ---- continuous-batching server -------------------------------------------
 * No counterpart in the reference.  Graphs are added once and stay resident
 * in HBM; requests (graph id, pattern) are submitted from any thread and
 * served by a scheduler thread that keeps `slots` batches in flight, each
 * with its own streams and workspace: whenever a slot frees, every pending
 * request (up to max_batch, FIFO) is admitted into one batch (the
 * mgpm_batch_run machinery over `lanes` streams), so requests that arrive
 * while others run join the next batch and batches on different slots
 * overlap on the device.  Admission is by HBM footprint: a graph whose
 * resident footprint would exceed the budget is refused (MGPM_ERR_CAPACITY),
 * and slot workspaces are charged against the same budget: a free slot
 * takes pending requests in arrival order while their workspaces fit, and
 * the first that does not fit waits for memory to free (batches in flight
 * to finish, idle workspaces to be released); only a request that cannot
 * fit even then fails, alone (MGPM_ERR_CAPACITY).  slots = 1, lanes = 1,
 * max_batch = 1 gives one launch per request in arrival order. */
typedef struct mgpm_server_config {
  int32_t device;
  int32_t slots;       /* batches in flight (0 = 3) */
  int32_t max_batch;   /* requests per batch (0 = 1024) */
  int32_t lanes;       /* streams per slot (0 = 2) */
  uint64_t hbm_budget; /* bytes for graphs + workspaces (0 = 90% of free) */
  int32_t pattern_census; /* as in mgpm_batch_opts */
  int32_t reserved32;
  uint64_t reserved[1];
} mgpm_server_config;

typedef struct mgpm_request {
  int32_t graph;       /* in: id from mgpm_server_add_graph */
  int32_t pattern;     /* in: MGPM_TC or MGPM_P1..MGPM_P8 */
  uint64_t arrival_us; /* in (replay only): offset from the replay start */
  uint64_t count;      /* out */
  int32_t status;      /* out */
  int32_t batch_size;  /* out: requests served by the same launch */
  double latency_ms;   /* out: completion - arrival */
} mgpm_request;

typedef struct mgpm_server_stats {
  uint64_t submitted, completed, failed, launches;
  uint64_t resident_bytes, workspace_bytes, budget_bytes;
  double busy_ms; /* wall time with at least one launch in flight */
  uint64_t deferred;        /* times a request waited for memory to free
                               (the intake stopped at it) */
  uint64_t deferred_failed; /* requests that could not fit even with every
                               idle workspace released (MGPM_ERR_CAPACITY) */
} mgpm_server_stats;

int mgpm_server_create(const mgpm_server_config *cfg, mgpm_server **out);
int mgpm_server_add_graph(mgpm_server *s, const mgpm_graph *g,
                          int32_t *graph_id);
int mgpm_server_submit(mgpm_server *s, int32_t graph_id, int32_t pattern,
                       uint64_t *ticket);
/* Blocks until the request finishes (timeout_ms < 0: no limit); fills
 * count/status/batch_size/latency_ms of *out (graph/pattern too). */
int mgpm_server_wait(mgpm_server *s, uint64_t ticket, int32_t timeout_ms,
                     mgpm_request *out);
/* Submits reqs[i] at reqs[i].arrival_us after the call starts (open-loop,
 * from the calling thread), then waits for all and fills the outputs. */
int mgpm_server_replay(mgpm_server *s, mgpm_request *reqs, int n);
int mgpm_server_stats_get(mgpm_server *s, mgpm_server_stats *st);
int mgpm_server_destroy(mgpm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* MGPM_H */

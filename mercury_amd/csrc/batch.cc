// C ABI: batched multi-graph counting (no counterpart in the reference,
// which runs one graph and one pattern per process), and the multi-query
// launcher it shares with the server.
#include "batch_exec.hpp"
#include "census.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <unordered_map>

using namespace mgpm;

namespace mgpm {

namespace {

// Pattern scratch per lane: 8 * cap ints per wave, grid sized to fill the
// device but kept under this many bytes.
constexpr uint64_t kLaneSlabBudget = uint64_t(2) << 30;
constexpr int kPatternLanes = 4;
// Pair-pass arrays per lane (the census workspace's budget).
constexpr size_t kLaneCensusBudget = size_t(4) << 30;
// Per query in the scratch buffer: 2 u64 counters and 4 u32 queue heads
// (zeroed by the fill kernel), and separately the census accumulators.
constexpr size_t kQueryHead = 32;
constexpr size_t kQueryAcc = kCensusSlots * sizeof(unsigned long long);

uint64_t slab_bytes(int64_t cap, int grid) {
  return uint64_t(pattern_waves_per_block()) * 8 * uint64_t(cap) *
         sizeof(int32_t) * uint64_t(grid);
}

}  // namespace

uint64_t query_cost(const mgpm_dgraph *g, int pattern) {
  if (pattern == MGPM_TC) return g->tc.probes + g->tc.m_plus + 1;
  // The pattern plan's task costs bound the pattern's pair tests.
  const PatternPlanResident &pr = g->pat.plan[pattern];
  uint64_t c = 0;
  for (uint64_t x : pr.edge_group_cost) c += x;
  for (uint64_t x : pr.vert_group_cost) c += x;
  return c + 1;
}

namespace {

// Census work relative to the triangle count's: the passes a census runs
// over the same DAG (triangles and sums; 4-cycles; 4- and 5-cliques; the
// pair passes).
uint64_t census_cost(const mgpm_dgraph *g, int passes) {
  const uint64_t tri = g->tc.probes + g->tc.m_plus + 1;
  if (passes & kCensusPassPairs) return 24 * tri;
  return tri * (2 + ((passes & kCensusPassC4) ? 1 : 0) +
                ((passes & kCensusPassK5) ? 6 : 0));
}

// The passes a lane's workspace is sized for per census query: P2, P3, P5
// and P6 need every pass; the others at most the 4-cycle and 5-clique
// passes, whatever other patterns of the same graph join them in a job.
int need_passes(int pattern) {
  const int all = census_passes(MGPM_P2);
  return (census_passes(pattern) & kCensusPassPairs)
             ? all
             : (kCensusPassC4 | kCensusPassK5);
}

}  // namespace

BatchExec::~BatchExec() {
  if (lanes_.empty() && !fork_ && !pinned_) return;
  (void)hipSetDevice(device_);
  for (hipStream_t s : lanes_) (void)hipStreamSynchronize(s);
  for (hipStream_t s : lanes_) (void)hipStreamDestroy(s);
  for (hipEvent_t e : lane_done_) (void)hipEventDestroy(e);
  if (fork_) (void)hipEventDestroy(fork_);
  if (pinned_) (void)hipHostFree(pinned_);
}

int BatchExec::init(int device, int cus, int lanes, bool census) {
  device_ = device;
  cus_ = cus;
  census_ = census;
  lanes = std::max(1, std::min(lanes, 16));
  MGPM_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
  for (int l = 0; l < lanes; ++l) {
    hipStream_t s = nullptr;
    MGPM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    lanes_.push_back(s);
    hipEvent_t e = nullptr;
    MGPM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    lane_done_.push_back(e);
  }
  slab_.reset(new DevBuf[static_cast<size_t>(lanes)]);
  census_work_.reset(new CensusWork[static_cast<size_t>(lanes)]);
  for (int l = 0; l < lanes; ++l) census_work_[l].budget = kLaneCensusBudget;
  slab_cap_.assign(static_cast<size_t>(lanes), 0);
  slab_grid_.assign(static_cast<size_t>(lanes), 0);
  order_.assign(static_cast<size_t>(lanes), {});
  return MGPM_OK;
}

uint64_t BatchExec::footprint() const {
  uint64_t b = scratch_.bytes;
  for (size_t l = 0; l < lanes_.size(); ++l)
    b += slab_[l].bytes + census_work_[l].bytes();
  return b;
}

void BatchExec::query_need(const mgpm_query &q, uint64_t *want_census,
                           uint64_t *want_slab) const {
  const mgpm_dgraph *g = q.graph;
  const int p = q.pattern;
  if (!g || p <= MGPM_TC || p > MGPM_P8) return;
  if (census_) {
    *want_census = std::max(
        *want_census, census_work_bytes(g, need_passes(p), kLaneCensusBudget));
  } else {
    const int64_t cap = std::max<int64_t>(g->host->max_degree(), 64);
    const uint64_t per_block = slab_bytes(cap, 1);
    const uint64_t grid = std::max<uint64_t>(
        1, std::min<uint64_t>(uint64_t(cus_) * 4, kLaneSlabBudget / per_block));
    *want_slab = std::max(*want_slab, per_block * grid);
  }
}

uint64_t BatchExec::growth(const mgpm_query *q, int nq) const {
  uint64_t want_census = 0, want_slab = 0;
  for (int i = 0; i < nq; ++i) query_need(q[i], &want_census, &want_slab);
  return growth_for(want_census, want_slab, nq);
}

void BatchExec::release() {
  (void)hipSetDevice(device_);
  for (hipStream_t s : lanes_) (void)hipStreamSynchronize(s);
  for (size_t l = 0; l < lanes_.size(); ++l) {
    slab_[l].reset();
    slab_cap_[l] = 0;
    slab_grid_[l] = 0;
    census_work_[l].~CensusWork();
    new (&census_work_[l]) CensusWork();
    census_work_[l].budget = kLaneCensusBudget;
  }
}

uint64_t BatchExec::growth_for(uint64_t want_census, uint64_t want_slab,
                               int nq) const {
  // A bound: every lane grown to the largest query's workspace.
  uint64_t g = 0;
  const int use = std::min(lanes(), std::max(nq, 1));
  for (int l = 0; l < use; ++l) {
    const uint64_t hc = census_work_[l].bytes(), hs = slab_[l].bytes;
    if (want_census > hc) g += want_census - hc;
    if (want_slab > hs) g += want_slab - hs;
  }
  const uint64_t sc = (kQueryHead + kQueryAcc) * uint64_t(std::max(nq, 64));
  if (sc > scratch_.bytes) g += sc - scratch_.bytes;
  return g;
}

int BatchExec::plan(const mgpm_query *q, int nq, uint64_t max_bytes) {
  const int L = lanes();
  for (int i = 0; i < nq; ++i) {
    if (!q[i].graph) return fail(MGPM_ERR_INVALID, "query without a graph");
    if (q[i].graph->device != device_)
      return fail(MGPM_ERR_INVALID, "query graph lives on another device");
    if (q[i].pattern < MGPM_TC || q[i].pattern > MGPM_P8)
      return fail(MGPM_ERR_PATTERN, "Not support.");
  }
  // Checked before anything is allocated or built.
  if (growth(q, nq) > max_bytes)
    return fail(MGPM_ERR_CAPACITY, "batch workspace exceeds the HBM budget");
  queries_.assign(static_cast<size_t>(nq), Planned{});
  jobs_.clear();
  // Census queries of one graph share one census job: the union of their
  // passes runs once and every query combines its count from it.
  std::unordered_map<const mgpm_dgraph *, int> job_of;
  for (int i = 0; i < nq; ++i) {
    mgpm_dgraph *g = q[i].graph;
    const int p = q[i].pattern;
    Planned pq{g, p, 0, -1};
    if (p == MGPM_TC) {
      pq.cost = query_cost(g, p);
    } else if (census_) {
      auto it = job_of.find(g);
      if (it == job_of.end()) {
        it = job_of.emplace(g, static_cast<int>(jobs_.size())).first;
        jobs_.push_back(Job{g, 0, 0});
      }
      pq.job = it->second;
      jobs_[pq.job].passes |= census_passes(p);
    } else {
      const int rc = pattern_make_resident(g, nullptr, p);
      if (rc) return rc;
      pq.cost = query_cost(g, p);
    }
    queries_[i] = pq;
  }
  for (Job &j : jobs_) j.cost = census_cost(j.g, j.passes);
  // Units of lane work: triangle and per-edge pattern queries (index i)
  // and census jobs (index -1 - j).
  struct Unit {
    int id;
    uint64_t cost;
    bool heavy_pattern;
  };
  std::vector<Unit> units;
  for (int i = 0; i < nq; ++i)
    if (queries_[i].job < 0)
      units.push_back({i, queries_[i].cost, queries_[i].pattern != MGPM_TC});
  for (int j = 0; j < static_cast<int>(jobs_.size()); ++j)
    units.push_back({-1 - j, jobs_[j].cost, false});
  // Pattern-heavy batches of per-edge kernels run on at most kPatternLanes
  // lanes: their kernels are throughput-bound with large working sets, and
  // more concurrent ones measured slower (TC + P2 + P4 on 512 graphs: 4
  // lanes 192 queries/s, 8 lanes 170, 16 lanes 151); triangle counts and
  // census jobs gain from every lane.
  uint64_t pat_cost = 0, other_cost = 0;
  for (const Unit &u : units) (u.heavy_pattern ? pat_cost : other_cost) += u.cost;
  const int use = pat_cost > other_cost ? std::min(L, kPatternLanes) : L;
  // Longest processing time first, each unit to the least-loaded lane.
  std::stable_sort(units.begin(), units.end(),
                   [](const Unit &a, const Unit &b) { return a.cost > b.cost; });
  std::vector<uint64_t> load(static_cast<size_t>(use), 0);
  for (auto &o : order_) o.clear();
  std::vector<int64_t> need_cap(static_cast<size_t>(L), 0);
  for (const Unit &u : units) {
    const int l = static_cast<int>(
        std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += u.cost;
    order_[l].push_back(u.id);
    if (u.id >= 0 && u.heavy_pattern)
      need_cap[l] = std::max<int64_t>(need_cap[l], queries_[u.id].g->pat.cap);
  }
  // Census jobs: the graph's DAG (built once, kept with the graph) and the
  // lane's workspace, grown to the largest job on the lane.
  for (int l = 0; l < L; ++l)
    for (int u : order_[l])
      if (u < 0) {
        const Job &j = jobs_[-1 - u];
        const int rc = census_prepare(j.g, j.passes, census_work_[l], lanes_[l]);
        if (rc) return rc;
      }
  for (int l = 0; l < L; ++l) {
    if (need_cap[l] == 0 || (slab_[l].p && slab_cap_[l] >= need_cap[l]))
      continue;
    const uint64_t per_block = slab_bytes(need_cap[l], 1);
    const int grid = static_cast<int>(std::max<uint64_t>(
        1, std::min<uint64_t>(uint64_t(cus_) * 4, kLaneSlabBudget / per_block)));
    MGPM_HIP(slab_[l].alloc(per_block * uint64_t(grid)));
    slab_cap_[l] = need_cap[l];
    slab_grid_[l] = grid;
  }
  const size_t per_query = kQueryHead + kQueryAcc;
  if (scratch_.bytes < per_query * size_t(nq))
    MGPM_HIP(scratch_.alloc(per_query * size_t(std::max(nq, 64))));
  if (pinned_n_ < static_cast<size_t>(nq)) {
    if (pinned_) (void)hipHostFree(pinned_);
    pinned_ = nullptr;
    pinned_n_ = 0;
    const size_t n = static_cast<size_t>(std::max(nq, 64));
    MGPM_HIP(hipHostMalloc(reinterpret_cast<void **>(&pinned_),
                           per_query * n, hipHostMallocDefault));
    pinned_n_ = n;
  }
  return MGPM_OK;
}

int BatchExec::enqueue(hipStream_t s, bool copy_back) {
  const int nq = size();
  if (nq == 0) return MGPM_OK;
  auto *counters = scratch_.as<unsigned long long>();
  auto *queues = reinterpret_cast<unsigned *>(counters + 2 * nq);
  auto *accs = reinterpret_cast<unsigned long long *>(scratch_.as<char>() +
                                                      kQueryHead * nq);
  if (device_zero(scratch_.p, kQueryHead / 4 * static_cast<size_t>(nq), s))
    return fail(MGPM_ERR_HIP, "batch reset launch failed");
  MGPM_HIP(hipEventRecord(fork_, s));
  for (int l = 0; l < lanes(); ++l) {
    if (order_[l].empty()) continue;
    hipStream_t ls = lanes_[l];
    MGPM_HIP(hipStreamWaitEvent(ls, fork_, 0));
    for (int u : order_[l]) {
      if (u < 0) {
        // The lane's jobs run in order on its one census workspace; the
        // accumulators are copied out before the next job reuses it.
        const int j = -1 - u;
        int rc = census_enqueue(jobs_[j].g, jobs_[j].passes, census_work_[l],
                                ls);
        if (rc) return rc;
        MGPM_HIP(hipMemcpyAsync(accs + kCensusSlots * size_t(j),
                                census_acc(census_work_[l]), kQueryAcc,
                                hipMemcpyDeviceToDevice, ls));
        continue;
      }
      const int i = u;
      const Planned &qp = queries_[i];
      if (qp.pattern == MGPM_TC) {
        TcDevicePlan p = qp.g->tc.dev;
        p.queues = queues + 4 * i;
        p.result = counters + 2 * i;
        TcLaunch tl;
        tl.cus = cus_;
        tl.items = qp.g->tc.items;
        tl.grid_cap = qp.g->tc.grid_cap;
        tl.cursor_min = qp.g->tc.cursor_min;
        tl.reset = false;
        if (tc_device_launch(p, tl, ls))
          return fail(MGPM_ERR_HIP, "triangle kernel launch failed");
      } else {
        PatternDevicePlan p = pattern_device_plan(qp.g, qp.pattern);
        p.slab = slab_[l].as<int32_t>();
        p.cap = slab_cap_[l];
        p.queues = queues + 4 * i;
        p.counters = counters + 2 * i;
        int grid = slab_grid_[l];
        if (qp.g->pat.grid_cap > 0) grid = std::min(grid, qp.g->pat.grid_cap);
        if (pattern_device_launch(p, qp.pattern, grid, ls, false))
          return fail(MGPM_ERR_HIP, "pattern kernel launch failed");
      }
    }
    MGPM_HIP(hipEventRecord(lane_done_[l], ls));
    MGPM_HIP(hipStreamWaitEvent(s, lane_done_[l], 0));
  }
  return copy_back ? enqueue_copy(s) : MGPM_OK;
}

int BatchExec::enqueue_copy(hipStream_t s) {
  const int nq = size();
  if (nq == 0) return MGPM_OK;
  // Counters, and the census accumulators of the jobs.
  const size_t bytes =
      kQueryHead * size_t(nq) + kQueryAcc * jobs_.size();
  MGPM_HIP(hipMemcpyAsync(pinned_, scratch_.p, bytes, hipMemcpyDeviceToHost,
                          s));
  return MGPM_OK;
}

int BatchExec::result(int i, uint64_t *count) const {
  const Planned &qp = queries_[i];
  const int nq = size();
  if (qp.job >= 0) {
    const unsigned long long *acc = reinterpret_cast<const unsigned long long *>(
        reinterpret_cast<const char *>(pinned_) + kQueryHead * nq +
        kQueryAcc * size_t(qp.job));
    return census_combine(qp.pattern, acc, count);
  }
  const uint64_t c1 = pinned_[2 * i], c2 = pinned_[2 * i + 1];
  // sm_multigpu.cu:209-232 on one rank: P4 and P6 report counters[1], the
  // other patterns counters[1] - counters[2]; TC is the single counter.
  const int p = qp.pattern;
  *count = (p == MGPM_TC || p == MGPM_P4 || p == MGPM_P6) ? c1 : c1 - c2;
  return MGPM_OK;
}

}  // namespace mgpm

struct mgpm_batch {
  int device = 0;
  int cus = 0;
  mgpm_batch_opts opts{};
  BatchExec exec;
  // The planned query set: (graph, pattern, TC plan generation).
  struct Key {
    const mgpm_dgraph *g;
    int pattern;
    uint64_t gen;
    bool operator==(const Key &o) const {
      return g == o.g && pattern == o.pattern && gen == o.gen;
    }
  };
  std::vector<Key> key;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Recorded launch sequence (opts.use_graph).
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

namespace {

void drop_graph(mgpm_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

std::vector<mgpm_batch::Key> key_of(const mgpm_query *q, int nq) {
  std::vector<mgpm_batch::Key> k(static_cast<size_t>(nq));
  for (int i = 0; i < nq; ++i)
    k[i] = {q[i].graph, q[i].pattern, q[i].graph ? q[i].graph->tc.generation : 0};
  return k;
}

// Records the whole launch sequence (fill, lane kernels, join) into a HIP
// graph.  The counter copy stays outside it.
int record(mgpm_batch *b) {
  drop_graph(b);
  MGPM_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  const int rc = b->exec.enqueue(b->cap_stream, false);
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->cap_stream, &graph);
  if (rc || ee != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : hip_fail(ee, "hipStreamEndCapture");
  }
  b->graph = graph;
  MGPM_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  return MGPM_OK;
}

}  // namespace

extern "C" {

int mgpm_batch_create(int device, const mgpm_batch_opts *opts,
                      mgpm_batch **out) {
  if (!out) return fail(MGPM_ERR_INVALID, "null argument");
  try {
    int rc = set_device(device);
    if (rc) return rc;
    std::unique_ptr<mgpm_batch> b(new mgpm_batch);
    b->device = device;
    if (opts) b->opts = *opts;
    if (b->opts.lanes <= 0) b->opts.lanes = 4;
    MGPM_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount,
                                   device));
    if ((rc = b->exec.init(device, b->cus, b->opts.lanes,
                           b->opts.pattern_census >= 0)))
      return rc;
    MGPM_HIP(hipEventCreate(&b->ev0));
    MGPM_HIP(hipEventCreate(&b->ev1));
    if (b->opts.use_graph)
      MGPM_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    *out = b.release();
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  }
}

int mgpm_batch_run(mgpm_batch *b, mgpm_query *queries, int nq, void *stream,
                   mgpm_stats *stats) {
  if (!b || (!queries && nq > 0) || nq < 0)
    return fail(MGPM_ERR_INVALID, "bad batch arguments");
  try {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = set_device(b->device);
    if (rc) return rc;
    for (int i = 0; i < nq; ++i) {
      queries[i].status = MGPM_ERR_FAILED;
      queries[i].count = 0;
    }
    if (nq == 0) return MGPM_OK;
    std::vector<mgpm_batch::Key> k = key_of(queries, nq);
    double prep = 0;
    if (k != b->key) {
      const auto tp = std::chrono::steady_clock::now();
      if (b->opts.autotune) {
        std::vector<mgpm_dgraph *> tc;
        for (int i = 0; i < nq; ++i)
          if (queries[i].pattern == MGPM_TC && queries[i].graph &&
              queries[i].graph->device == b->device)
            tc.push_back(queries[i].graph);
        if ((rc = autotune_size_classes(tc, 3))) return rc;
        k = key_of(queries, nq);
      }
      b->key.clear();
      drop_graph(b);
      if ((rc = b->exec.plan(queries, nq, UINT64_MAX))) return rc;
      if (b->opts.use_graph && (rc = record(b))) return rc;
      b->key = key_of(queries, nq);  // after any plan built on the way
      prep = ms_since(tp);
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    // A recorded graph replays on the caller's stream, or on the batch's
    // own stream when the caller passes the default stream (the call is
    // synchronous either way).
    if (b->gexec && !s) s = b->cap_stream;
    MGPM_HIP(hipEventRecord(b->ev0, s));
    if (b->gexec) {
      MGPM_HIP(hipGraphLaunch(b->gexec, s));
      if ((rc = b->exec.enqueue_copy(s))) return rc;
    } else if ((rc = b->exec.enqueue(s))) {
      return rc;
    }
    MGPM_HIP(hipEventRecord(b->ev1, s));
    MGPM_HIP(hipStreamSynchronize(s));
    float kms = 0;
    MGPM_HIP(hipEventElapsedTime(&kms, b->ev0, b->ev1));
    int worst = MGPM_OK;
    for (int i = 0; i < nq; ++i) {
      uint64_t c = 0;
      queries[i].status = b->exec.result(i, &c);
      queries[i].count = c;
      if (queries[i].status) worst = queries[i].status;
    }
    if (worst) return fail(worst, "batch: a census self-check failed");
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->kernel_ms = kms;
      stats->prep_ms = prep;
      stats->path = b->gexec ? 101 : 100;
      stats->e2e_ms = ms_since(t0);
    }
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::exception &ex) {
    return fail(MGPM_ERR_INTERNAL, ex.what());
  }
}

int mgpm_batch_footprint(const mgpm_batch *b, uint64_t *bytes) {
  if (!b || !bytes) return fail(MGPM_ERR_INVALID, "null argument");
  *bytes = b->exec.footprint();
  return MGPM_OK;
}

int mgpm_batch_free(mgpm_batch *b) {
  if (!b) return MGPM_OK;
  (void)hipSetDevice(b->device);
  if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
  drop_graph(b);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
  return MGPM_OK;
}

}  // extern "C"

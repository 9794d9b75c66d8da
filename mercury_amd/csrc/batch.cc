// C ABI: batched multi-graph counting (no counterpart in the reference,
// which runs one graph and one pattern per process), and the multi-query
// launcher it shares with the server.
#include "batch_exec.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

using namespace mgpm;

namespace mgpm {

namespace {

// Pattern scratch per lane: 8 * cap ints per wave, grid sized to fill the
// device but kept under this many bytes.
constexpr uint64_t kLaneSlabBudget = uint64_t(2) << 30;
constexpr int kPatternLanes = 4;

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

BatchExec::~BatchExec() {
  if (lanes_.empty() && !fork_ && !pinned_) return;
  (void)hipSetDevice(device_);
  for (hipStream_t s : lanes_) (void)hipStreamSynchronize(s);
  for (hipStream_t s : lanes_) (void)hipStreamDestroy(s);
  for (hipEvent_t e : lane_done_) (void)hipEventDestroy(e);
  if (fork_) (void)hipEventDestroy(fork_);
  if (pinned_) (void)hipHostFree(pinned_);
}

int BatchExec::init(int device, int cus, int lanes) {
  device_ = device;
  cus_ = cus;
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
  slab_cap_.assign(static_cast<size_t>(lanes), 0);
  slab_grid_.assign(static_cast<size_t>(lanes), 0);
  order_.assign(static_cast<size_t>(lanes), {});
  return MGPM_OK;
}

uint64_t BatchExec::footprint() const {
  uint64_t b = scratch_.bytes;
  for (size_t l = 0; l < lanes_.size(); ++l) b += slab_[l].bytes;
  return b;
}

int BatchExec::plan(const mgpm_query *q, int nq, uint64_t max_bytes) {
  const int L = lanes();
  queries_.assign(static_cast<size_t>(nq), Planned{});
  for (int i = 0; i < nq; ++i) {
    mgpm_dgraph *g = q[i].graph;
    const int p = q[i].pattern;
    if (!g) return fail(MGPM_ERR_INVALID, "query without a graph");
    if (g->device != device_)
      return fail(MGPM_ERR_INVALID, "query graph lives on another device");
    if (p < MGPM_TC || p > MGPM_P8) return fail(MGPM_ERR_PATTERN, "Not support.");
    if (p != MGPM_TC) {
      const int rc = pattern_make_resident(g, nullptr, p);
      if (rc) return rc;
    }
    queries_[i] = Planned{g, p, query_cost(g, p)};
  }
  // Pattern-heavy batches run on at most kPatternLanes lanes: their kernels
  // are throughput-bound with large working sets, and more concurrent ones
  // measured slower (TC + P2 + P4 on 512 graphs: 4 lanes 192 queries/s,
  // 8 lanes 170, 16 lanes 151); triangle counts gain from every lane.
  uint64_t pat_cost = 0, tc_cost = 0;
  for (const Planned &pq : queries_)
    (pq.pattern == MGPM_TC ? tc_cost : pat_cost) += pq.cost;
  const int use = pat_cost > tc_cost ? std::min(L, kPatternLanes) : L;
  // Longest processing time first, each query to the least-loaded lane.
  std::vector<int> idx(static_cast<size_t>(nq));
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return queries_[a].cost > queries_[b].cost;
  });
  std::vector<uint64_t> load(static_cast<size_t>(use), 0);
  for (auto &o : order_) o.clear();
  std::vector<int64_t> need_cap(static_cast<size_t>(L), 0);
  for (int i : idx) {
    const int l = static_cast<int>(
        std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += queries_[i].cost;
    order_[l].push_back(i);
    if (queries_[i].pattern != MGPM_TC)
      need_cap[l] = std::max<int64_t>(need_cap[l], queries_[i].g->pat.cap);
  }
  // Workspace sizes; grow-only, checked against the limit first.
  std::vector<int> grid(static_cast<size_t>(L), 0);
  std::vector<uint64_t> want(static_cast<size_t>(L), 0);
  uint64_t total = std::max<uint64_t>(scratch_.bytes, 32ull * uint64_t(nq));
  for (int l = 0; l < L; ++l) {
    if (need_cap[l] == 0 ||
        (slab_[l].p && slab_cap_[l] >= need_cap[l])) {
      total += slab_[l].bytes;
      continue;
    }
    const uint64_t per_block = slab_bytes(need_cap[l], 1);
    grid[l] = static_cast<int>(std::max<uint64_t>(
        1, std::min<uint64_t>(uint64_t(cus_) * 4, kLaneSlabBudget / per_block)));
    want[l] = per_block * uint64_t(grid[l]);
    total += want[l];
  }
  if (total > max_bytes)
    return fail(MGPM_ERR_CAPACITY, "batch workspace exceeds the HBM budget");
  for (int l = 0; l < L; ++l) {
    if (!want[l]) continue;
    MGPM_HIP(slab_[l].alloc(want[l]));
    slab_cap_[l] = need_cap[l];
    slab_grid_[l] = grid[l];
  }
  if (scratch_.bytes < 32ull * uint64_t(nq))
    MGPM_HIP(scratch_.alloc(32ull * uint64_t(std::max(nq, 64))));
  if (pinned_n_ < static_cast<size_t>(nq)) {
    if (pinned_) (void)hipHostFree(pinned_);
    pinned_ = nullptr;
    pinned_n_ = 0;
    const size_t n = static_cast<size_t>(std::max(nq, 64));
    MGPM_HIP(hipHostMalloc(reinterpret_cast<void **>(&pinned_),
                           2 * n * sizeof(unsigned long long),
                           hipHostMallocDefault));
    pinned_n_ = n;
  }
  return MGPM_OK;
}

int BatchExec::enqueue(hipStream_t s, bool copy_back) {
  const int nq = size();
  if (nq == 0) return MGPM_OK;
  auto *counters = scratch_.as<unsigned long long>();
  auto *queues = reinterpret_cast<unsigned *>(counters + 2 * nq);
  if (device_zero(scratch_.p, 8 * static_cast<size_t>(nq), s))
    return fail(MGPM_ERR_HIP, "batch reset launch failed");
  MGPM_HIP(hipEventRecord(fork_, s));
  for (int l = 0; l < lanes(); ++l) {
    if (order_[l].empty()) continue;
    hipStream_t ls = lanes_[l];
    MGPM_HIP(hipStreamWaitEvent(ls, fork_, 0));
    for (int i : order_[l]) {
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
  MGPM_HIP(hipMemcpyAsync(pinned_, scratch_.p,
                          2 * static_cast<size_t>(nq) * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
  return MGPM_OK;
}

uint64_t BatchExec::count(int i) const {
  const uint64_t c1 = pinned_[2 * i], c2 = pinned_[2 * i + 1];
  const int p = queries_[i].pattern;
  // sm_multigpu.cu:209-232 on one rank: P4 and P6 report counters[1], the
  // other patterns counters[1] - counters[2]; TC is the single counter.
  return (p == MGPM_TC || p == MGPM_P4 || p == MGPM_P6) ? c1 : c1 - c2;
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
    if ((rc = b->exec.init(device, b->cus, b->opts.lanes))) return rc;
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
    for (int i = 0; i < nq; ++i) {
      queries[i].count = b->exec.count(i);
      queries[i].status = MGPM_OK;
    }
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

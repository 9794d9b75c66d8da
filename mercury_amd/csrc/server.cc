// C ABI: continuous-batching count server (no counterpart in the
// reference).  One scheduler thread; `slots` batches in flight, each on its
// own streams with its own workspace; pending requests are admitted into
// one batch whenever a slot frees.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "batch_exec.hpp"

using namespace mgpm;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  int32_t graph = 0, pattern = 0;
  Clock::time_point arrival;
  uint64_t count = 0;
  int32_t status = MGPM_ERR_FAILED;
  int32_t batch_size = 0;
  double latency_ms = 0;
  bool done = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  BatchExec exec;
  std::vector<uint64_t> tickets;
  bool busy = false;
};

}  // namespace

struct mgpm_server {
  mgpm_server_config cfg{};
  int cus = 0;
  uint64_t budget = 0, resident = 0, workspace = 0;
  std::vector<mgpm_dgraph *> graphs;
  std::vector<uint64_t> graph_bytes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, Req> reqs;
  uint64_t next_ticket = 1;
  std::deque<Slot> slots;  // stable addresses, never resized after create
  std::thread sched;
  std::atomic<bool> stop{false};
  mgpm_server_stats st{};
  Clock::time_point busy_since;
  int inflight = 0;
  int waiters = 0;       // threads inside mgpm_server_wait
  bool closing = false;  // mgpm_server_destroy has failed what was left
};

namespace {

void finish(mgpm_server *s, uint64_t t, uint64_t count, int status,
            int batch, Clock::time_point now);

mgpm_query query_of(const mgpm_server *s, const Req &r) {
  mgpm_query q{};
  q.graph = s->graphs[r.graph];
  q.pattern = r.pattern;
  return q;
}

// Launches the requests `tickets` on slot `sl` (server lock held); their
// workspace growth was checked against the budget by the caller.
int launch_slot(mgpm_server *s, Slot &sl, const std::vector<uint64_t> &tickets) {
  std::vector<mgpm_query> q(tickets.size());
  for (size_t i = 0; i < tickets.size(); ++i)
    q[i] = query_of(s, s->reqs[tickets[i]]);
  const uint64_t before = sl.exec.footprint();
  const uint64_t others = s->resident + s->workspace - before;
  const uint64_t room = s->budget > others + before
                            ? s->budget - others - before
                            : 0;
  int rc = sl.exec.plan(q.data(), static_cast<int>(q.size()), room);
  s->workspace = s->workspace - before + sl.exec.footprint();
  if (rc) return rc;
  if ((rc = sl.exec.enqueue(sl.stream))) return rc;
  MGPM_HIP(hipEventRecord(sl.done, sl.stream));
  sl.tickets = tickets;
  sl.busy = true;
  if (s->inflight++ == 0) s->busy_since = Clock::now();
  s->st.launches++;
  return MGPM_OK;
}

// Budget left for slot `sl`'s workspace to grow into.
uint64_t room_for(const mgpm_server *s) {
  const uint64_t used = s->resident + s->workspace;
  return s->budget > used ? s->budget - used : 0;
}

// Takes pending requests in arrival order for free slot `sl` while their
// workspaces fit the budget (server lock held).  The first request that
// does not fit stops the intake and waits: when no batch is in flight, the
// idle slots' workspaces are released for it, and if it still cannot fit
// it fails alone (MGPM_ERR_CAPACITY) while the rest keep their turn.
std::vector<uint64_t> admit(mgpm_server *s, Slot &sl) {
  std::vector<uint64_t> take;
  uint64_t wc = 0, ws = 0;
  const size_t cap = static_cast<size_t>(s->cfg.max_batch);
  while (!s->pending.empty() && take.size() < cap) {
    const uint64_t t = s->pending.front();
    const mgpm_query q = query_of(s, s->reqs[t]);
    uint64_t c2 = wc, s2 = ws;
    sl.exec.query_need(q, &c2, &s2);
    const int nq = static_cast<int>(take.size()) + 1;
    if (sl.exec.growth_for(c2, s2, nq) <= room_for(s)) {
      wc = c2;
      ws = s2;
      take.push_back(t);
      s->pending.pop_front();
      continue;
    }
    if (!take.empty() || s->inflight > 0) {  // wait for memory to free
      s->st.deferred++;
      break;
    }
    // Nothing in flight: free every idle slot's workspace and try again.
    uint64_t freed = 0;
    for (Slot &o : s->slots) {
      if (o.busy) continue;
      const uint64_t b = o.exec.footprint();
      o.exec.release();
      freed += b - o.exec.footprint();
    }
    s->workspace -= std::min(s->workspace, freed);
    if (sl.exec.growth_for(c2, s2, nq) <= room_for(s)) continue;
    s->pending.pop_front();  // can never fit: fails alone
    s->st.deferred_failed++;
    finish(s, t, 0, MGPM_ERR_CAPACITY, 0, Clock::now());
    s->cv_done.notify_all();
  }
  return take;
}

void finish(mgpm_server *s, uint64_t t, uint64_t count, int status,
            int batch, Clock::time_point now) {
  Req &r = s->reqs[t];
  r.count = count;
  r.status = status;
  r.batch_size = batch;
  r.latency_ms =
      std::chrono::duration<double, std::milli>(now - r.arrival).count();
  r.done = true;
  if (status == MGPM_OK) s->st.completed++;
  else s->st.failed++;
}

void scheduler(mgpm_server *s) {
  (void)hipSetDevice(s->cfg.device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop.load()) {
    bool progressed = false;
    // Completions.  A slot whose work ended in an error (a failed launch
    // or kernel reported by the event) finishes its requests with
    // MGPM_ERR_HIP instead of staying busy forever.
    for (Slot &sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      const auto now = Clock::now();
      const int nb = static_cast<int>(sl.tickets.size());
      if (q != hipSuccess) set_error(std::string("server batch: ") +
                                     hipGetErrorString(q));
      for (int i = 0; i < nb; ++i) {
        uint64_t c = 0;
        const int st = q == hipSuccess ? sl.exec.result(i, &c) : MGPM_ERR_HIP;
        finish(s, sl.tickets[i], c, st, nb, now);
      }
      sl.tickets.clear();
      sl.busy = false;
      if (--s->inflight == 0)
        s->st.busy_ms +=
            std::chrono::duration<double, std::milli>(now - s->busy_since)
                .count();
      progressed = true;
    }
    if (progressed) s->cv_done.notify_all();
    // Admissions: every free slot takes pending requests, FIFO, as far as
    // the memory budget allows.
    for (Slot &sl : s->slots) {
      if (sl.busy || s->pending.empty()) continue;
      std::vector<uint64_t> take = admit(s, sl);
      if (take.empty()) continue;
      const int rc = launch_slot(s, sl, take);
      if (rc) {  // isolate the failure to these requests; keep serving
        const auto now = Clock::now();
        for (uint64_t t : take) finish(s, t, 0, rc, 0, now);
        s->cv_done.notify_all();
      }
      progressed = true;
    }
    if (!progressed) {
      if (s->inflight == 0 && s->pending.empty())
        s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
      else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      }
    }
  }
}

}  // namespace

extern "C" {

int mgpm_server_create(const mgpm_server_config *cfg, mgpm_server **out) {
  if (!cfg || !out) return fail(MGPM_ERR_INVALID, "null argument");
  try {
    int rc = set_device(cfg->device);
    if (rc) return rc;
    std::unique_ptr<mgpm_server> s(new mgpm_server);
    s->cfg = *cfg;
    if (s->cfg.slots <= 0) s->cfg.slots = 3;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 1024;
    if (s->cfg.lanes <= 0) s->cfg.lanes = 2;
    MGPM_HIP(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount,
                                   cfg->device));
    size_t free_b = 0, total_b = 0;
    MGPM_HIP(hipMemGetInfo(&free_b, &total_b));
    s->budget = cfg->hbm_budget ? cfg->hbm_budget : free_b / 10 * 9;
    s->slots.resize(static_cast<size_t>(s->cfg.slots));
    for (Slot &sl : s->slots) {
      MGPM_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
      MGPM_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
      if ((rc = sl.exec.init(cfg->device, s->cus, s->cfg.lanes,
                             s->cfg.pattern_census >= 0)))
        return rc;
    }
    s->st.budget_bytes = s->budget;
    s->sched = std::thread(scheduler, s.get());
    *out = s.release();
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::system_error &e) {
    return fail(MGPM_ERR_INTERNAL, e.what());
  }
}

int mgpm_server_add_graph(mgpm_server *s, const mgpm_graph *g,
                          int32_t *graph_id) {
  if (!s || !g || !graph_id) return fail(MGPM_ERR_INVALID, "null argument");
  try {
    int rc = set_device(s->cfg.device);
    if (rc) return rc;
    // Upload (TC plan) and build what the pattern counts keep with the
    // graph -- the census DAG and pair-table layout, or every pattern's
    // per-edge plan -- then admit by footprint (the slots hold the
    // workspaces, so the graph holds none).
    mgpm_dgraph *dg = nullptr;
    if ((rc = mgpm_dgraph_upload(g, s->cfg.device, nullptr, &dg))) return rc;
    if (s->cfg.pattern_census >= 0) {
      rc = census_graph_prepare(dg, nullptr);
    } else {
      for (int p = MGPM_P1; p <= MGPM_P8 && !rc; ++p)
        rc = pattern_make_resident(dg, nullptr, p);
    }
    if (rc) {
      mgpm_dgraph_free(dg);
      return rc;
    }
    uint64_t bytes = 0;
    mgpm_dgraph_footprint(dg, &bytes);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->resident + s->workspace + bytes > s->budget) {
      mgpm_dgraph_free(dg);
      return fail(MGPM_ERR_CAPACITY, "graph does not fit the HBM budget");
    }
    s->resident += bytes;
    s->graphs.push_back(dg);
    s->graph_bytes.push_back(bytes);
    *graph_id = static_cast<int32_t>(s->graphs.size()) - 1;
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  }
}

int mgpm_server_submit(mgpm_server *s, int32_t graph_id, int32_t pattern,
                       uint64_t *ticket) {
  if (!s || !ticket) return fail(MGPM_ERR_INVALID, "null argument");
  if (pattern < MGPM_TC || pattern > MGPM_P8)
    return fail(MGPM_ERR_PATTERN, "Not support.");
  std::lock_guard<std::mutex> lk(s->mu);
  if (graph_id < 0 || graph_id >= static_cast<int32_t>(s->graphs.size()))
    return fail(MGPM_ERR_NOTFOUND, "unknown graph id");
  const uint64_t t = s->next_ticket++;
  Req &r = s->reqs[t];
  r.graph = graph_id;
  r.pattern = pattern;
  r.arrival = Clock::now();
  s->pending.push_back(t);
  s->st.submitted++;
  *ticket = t;
  s->cv_work.notify_one();
  return MGPM_OK;
}

int mgpm_server_wait(mgpm_server *s, uint64_t ticket, int32_t timeout_ms,
                     mgpm_request *out) {
  if (!s) return fail(MGPM_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(MGPM_ERR_NOTFOUND, "unknown ticket");
  auto ready = [&] { return s->reqs[ticket].done; };
  ++s->waiters;
  bool ok = true;
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                ready);
  if (--s->waiters == 0 && s->closing) s->cv_done.notify_all();
  if (!ok) return fail(MGPM_ERR_TIMEOUT, "request still running");
  const Req r = s->reqs[ticket];
  s->reqs.erase(ticket);
  if (out) {
    out->graph = r.graph;
    out->pattern = r.pattern;
    out->count = r.count;
    out->status = r.status;
    out->batch_size = r.batch_size;
    out->latency_ms = r.latency_ms;
  }
  return r.status;
}

int mgpm_server_replay(mgpm_server *s, mgpm_request *reqs, int n) {
  if (!s || (!reqs && n > 0) || n < 0)
    return fail(MGPM_ERR_INVALID, "bad replay arguments");
  std::vector<uint64_t> tickets(static_cast<size_t>(n), 0);
  std::vector<int> sub_rc(static_cast<size_t>(n), MGPM_OK);
  const auto t0 = Clock::now();
  for (int i = 0; i < n; ++i) {
    const auto at = t0 + std::chrono::microseconds(reqs[i].arrival_us);
    while (Clock::now() < at) {
      const auto left = at - Clock::now();
      if (left > std::chrono::microseconds(200))
        std::this_thread::sleep_for(left - std::chrono::microseconds(100));
    }
    sub_rc[i] = mgpm_server_submit(s, reqs[i].graph, reqs[i].pattern,
                                   &tickets[i]);
  }
  int worst = MGPM_OK;
  for (int i = 0; i < n; ++i) {
    if (sub_rc[i]) {
      reqs[i].status = sub_rc[i];
      worst = sub_rc[i];
      continue;
    }
    mgpm_request r;
    const int rc = mgpm_server_wait(s, tickets[i], -1, &r);
    reqs[i].count = r.count;
    reqs[i].status = rc;
    reqs[i].batch_size = r.batch_size;
    reqs[i].latency_ms = r.latency_ms;
    if (rc) worst = rc;
  }
  return worst == MGPM_OK ? MGPM_OK : fail(MGPM_ERR_FAILED,
                                           "some requests failed");
}

int mgpm_server_stats_get(mgpm_server *s, mgpm_server_stats *st) {
  if (!s || !st) return fail(MGPM_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  st->resident_bytes = s->resident;
  st->workspace_bytes = s->workspace;
  st->budget_bytes = s->budget;
  return MGPM_OK;
}

int mgpm_server_destroy(mgpm_server *s) {
  if (!s) return MGPM_OK;
  s->stop.store(true);
  s->cv_work.notify_all();
  if (s->sched.joinable()) s->sched.join();
  (void)hipSetDevice(s->cfg.device);
  for (Slot &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  {
    // Requests still pending or in flight fail; wake their waiters, and
    // wait until every waiter has left before the server goes away.
    std::unique_lock<std::mutex> lk(s->mu);
    const auto now = Clock::now();
    for (auto &kv : s->reqs)
      if (!kv.second.done) finish(s, kv.first, 0, MGPM_ERR_FAILED, 0, now);
    s->pending.clear();
    s->closing = true;
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  s->slots.clear();  // lane streams and workspaces
  for (mgpm_dgraph *dg : s->graphs) mgpm_dgraph_free(dg);
  delete s;
  return MGPM_OK;
}

}  // extern "C"

// C ABI: host graphs, device graphs and triangle counting.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>

#include "capi_common.hpp"

namespace mgpm {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

uint64_t next_generation() {
  static std::atomic<uint64_t> seq{0};
  return seq.fetch_add(1, std::memory_order_relaxed) + 1;
}

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return MGPM_ERR_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    return MGPM_ERR_NODEVICE;
  return MGPM_ERR_HIP;
}

size_t TcResident::bytes() const {
  size_t b = col.bytes + partners.bytes + laid.bytes + laid_off.bytes;
  for (int c = 0; c < kTcNumClasses; ++c)
    b += tasks[c].bytes + groups[c].bytes;
  return b;
}

TcPlanParams tc_params_from(const mgpm_tune *t, int64_t arcs) {
  TcPlanParams p;
  if (arcs > 0) p.chunk_work = tc_default_chunk(arcs);
  p.small_max = kTcSmallMaxDeg;
  p.wave_max = kTcWaveMaxDeg;
  p.block_max = kTcBlockMaxDeg;
  if (!t) return p;
  auto pick = [](int32_t v, int32_t def, int32_t cap) {
    if (v == 0) return def;
    if (v < 0) return 0;
    return std::min(v, cap);
  };
  p.small_max = pick(t->small_max_deg, kTcSmallMaxDeg, kTcSmallMaxDeg);
  p.wave_max = std::max(p.small_max,
                        pick(t->wave_max_deg, kTcWaveMaxDeg, kTcWaveMaxDeg));
  p.block_max = std::max(
      p.wave_max, pick(t->block_max_deg, kTcBlockMaxDeg, kTcBlockMaxDeg));
  if (t->chunk_work > 0) p.chunk_work = t->chunk_work;
  if (t->group_work > 0) p.group_work = t->group_work;
  return p;
}

namespace {

bool same_params(const TcPlanParams &a, const TcPlanParams &b) {
  return a.small_max == b.small_max && a.wave_max == b.wave_max &&
         a.block_max == b.block_max && a.chunk_work == b.chunk_work &&
         a.group_work == b.group_work;
}

template <class T>
int upload_vec(DevBuf &buf, const std::vector<T> &v, hipStream_t s) {
  MGPM_HIP(buf.alloc(v.size() * sizeof(T)));
  if (!v.empty())
    MGPM_HIP(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T),
                            hipMemcpyHostToDevice, s));
  return MGPM_OK;
}

}  // namespace

// (Re)builds the triangle plan on the host and uploads it.
int tc_make_resident(mgpm_dgraph *dg, const TcPlanParams &params,
                     hipStream_t s) {
  TcPlan plan;
  int rc = build_tc_plan(*dg->host, params, plan);
  if (rc == MGPM_ERR_CAPACITY)
    return fail(rc, "triangle plan: DAG has 2^32 or more arcs");
  if (rc) return fail(rc, "triangle plan: a list lies outside the DAG");
  // Upload into fresh buffers and install them only when every copy has
  // landed, so a failure leaves the resident plan untouched.
  DevBuf col, partners, tasks[kTcNumClasses], groups[kTcNumClasses];
  TcDevicePlan p;
  if ((rc = upload_vec(col, plan.dag_col, s))) return rc;
  if ((rc = upload_vec(partners, plan.partners, s))) return rc;
  for (int c = 0; c < kTcNumClasses; ++c) {
    if ((rc = upload_vec(tasks[c], plan.tasks[c], s))) return rc;
    if ((rc = upload_vec(groups[c], plan.group_off[c], s))) return rc;
    p.tasks[c] = tasks[c].as<TcTask>();
    p.ntasks[c] = static_cast<int>(plan.tasks[c].size());
    p.groups[c] = groups[c].as<int32_t>();
    p.ngroups[c] = static_cast<int>(plan.group_off[c].size()) - 1;
    p.class_max_tlen[c] = plan.class_max_tlen[c];
  }
  p.col = col.as<int32_t>();
  p.ncol = static_cast<uint32_t>(plan.m_plus);
  p.partners = partners.as<TcPartner>();
  p.queues = dg->scratch.as<unsigned>();
  p.result = reinterpret_cast<unsigned long long *>(dg->scratch.as<char>() + 64);
  // The host vectors die with `plan`: wait for the copies.
  MGPM_HIP(hipStreamSynchronize(s));
  TcResident &r = dg->tc;
  r.col.swap(col);
  r.partners.swap(partners);
  for (int c = 0; c < kTcNumClasses; ++c) {
    r.tasks[c].swap(tasks[c]);
    r.groups[c].swap(groups[c]);
  }
  r.dev = p;
  r.params = params;
  r.on_device = false;
  r.generation = next_generation();
  r.drop_layout();
  r.npart = plan.partners.size();
  r.s_plus = plan.s_plus;
  r.probes = plan.probes;
  r.m_plus = static_cast<uint64_t>(plan.m_plus);
  r.max_out = plan.max_out;
  if (!r.tuned) r.items = default_tc_items(plan.probes, plan.partners.size());
  return MGPM_OK;
}

int tc_replan(mgpm_dgraph *dg, const TcPlanParams &params, hipStream_t s) {
  return dg->tc.on_device ? tc_make_resident_device(dg, params, s)
                          : tc_make_resident(dg, params, s);
}

// Bytes of the laid-out partner lists of the resident plan.
uint64_t tc_layout_bytes(const TcResident &r) {
  return 4ull * r.probes + 8ull * (r.npart + 1);
}

// Builds the laid-out partner lists of the resident plan (tc_layout) if
// they are not there.
int tc_layout_build(mgpm_dgraph *dg, hipStream_t s) {
  TcResident &r = dg->tc;
  if (r.laid.p && r.laid_off.p) return MGPM_OK;
  DevBuf items, item_off, lens, tmp;
  const size_t tb = tc_layout_scan_bytes(r.npart);
  if (tb == 0) return fail(MGPM_ERR_HIP, "scan sizing failed");
  MGPM_HIP(items.alloc(std::max<uint64_t>(r.probes, 1) * 4));
  MGPM_HIP(item_off.alloc((r.npart + 1) * 8));
  MGPM_HIP(lens.alloc((r.npart + 1) * 8));
  MGPM_HIP(tmp.alloc(tb));
  if (tc_layout(r.dev.col, r.dev.partners, r.npart, lens.as<uint64_t>(),
                item_off.as<uint64_t>(), items.as<int32_t>(), tmp.p, tb,
                dg->cus, s))
    return fail(MGPM_ERR_HIP, "layout kernel launch failed");
  uint64_t total = 0;
  MGPM_HIP(hipMemcpyAsync(&total, item_off.as<uint64_t>() + r.npart, 8,
                          hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  if (total != r.probes)
    return fail(MGPM_ERR_INTERNAL, "laid-out lists: item count mismatch");
  r.laid.swap(items);
  r.laid_off.swap(item_off);
  return MGPM_OK;
}

namespace {

int wrap_graph(std::shared_ptr<CsrGraph> g, mgpm_graph **out) {
  mgpm_graph *h = new (std::nothrow) mgpm_graph;
  if (!h) return fail(MGPM_ERR_NOMEM, "out of host memory");
  h->g = std::move(g);
  *out = h;
  return MGPM_OK;
}

}  // namespace

int set_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(MGPM_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= count)
    return fail(MGPM_ERR_NODEVICE, "device index out of range");
  MGPM_HIP(hipSetDevice(device));
  return MGPM_OK;
}

const char *last_error_cstr() { return g_last_error.c_str(); }

}  // namespace mgpm

using namespace mgpm;

#define MGPM_GUARD_BEGIN try {
#define MGPM_GUARD_END                                            \
  }                                                               \
  catch (const std::bad_alloc &) {                                \
    return fail(MGPM_ERR_NOMEM, "out of host memory");            \
  }                                                               \
  catch (const std::exception &ex) {                              \
    return fail(MGPM_ERR_INTERNAL, ex.what());                    \
  }                                                               \
  catch (...) {                                                   \
    return fail(MGPM_ERR_INTERNAL, "unknown exception");          \
  }

extern "C" {

const char *mgpm_strerror(int code) {
  switch (code) {
    case MGPM_OK: return "ok";
    case MGPM_ERR_IO: return "i/o error";
    case MGPM_ERR_FORMAT: return "bad graph format";
    case MGPM_ERR_INVALID: return "invalid argument";
    case MGPM_ERR_NOMEM: return "out of memory";
    case MGPM_ERR_HIP: return "HIP runtime error";
    case MGPM_ERR_NODEVICE: return "no such device";
    case MGPM_ERR_GRAPH: return "graph is not simple, sorted and symmetric";
    case MGPM_ERR_PATTERN: return "Not support.";
    case MGPM_ERR_CAPACITY: return "capacity exceeded";
    case MGPM_ERR_BUSY: return "busy";
    case MGPM_ERR_NOTFOUND: return "not found";
    case MGPM_ERR_TIMEOUT: return "timed out";
    case MGPM_ERR_FAILED: return "request failed";
    case MGPM_ERR_INTERNAL: return "internal error";
    default: return "unknown error";
  }
}

int mgpm_abi_version(void) { return MGPM_ABI_VERSION; }

const char *mgpm_last_error(void) { return last_error_cstr(); }

// ------------------------------------------------------------ host graphs

int mgpm_graph_load(const char *prefix, mgpm_graph **out) {
  return mgpm_graph_load_ex(prefix, 0, out);
}

int mgpm_graph_load_ex(const char *prefix, int32_t flags, mgpm_graph **out) {
  if (!prefix || !out) return fail(MGPM_ERR_INVALID, "null argument");
  if (flags & ~(MGPM_LOAD_VLABELS | MGPM_LOAD_ELABELS))
    return fail(MGPM_ERR_INVALID, "unknown load flags");
  MGPM_GUARD_BEGIN
  auto g = std::make_shared<CsrGraph>();
  std::string err;
  int rc = read_graph(prefix, *g, err,
                      ((flags & MGPM_LOAD_VLABELS) ? kLoadVertexLabels : 0) |
                          ((flags & MGPM_LOAD_ELABELS) ? kLoadEdgeLabels : 0));
  if (rc) return fail(rc, err);
  return wrap_graph(std::move(g), out);
  MGPM_GUARD_END
}

int mgpm_graph_labels(const mgpm_graph *g, const uint8_t **vlabels,
                      int32_t *num_vertex_classes, const uint16_t **elabels,
                      int32_t *num_edge_classes) {
  if (!g) return fail(MGPM_ERR_INVALID, "null argument");
  const CsrGraph &c = *g->g;
  const bool hv = !c.vlabels.empty() || (c.n == 0 && c.num_vertex_classes);
  const bool he = !c.elabels.empty() || (c.nnz == 0 && c.num_edge_classes);
  if (vlabels) *vlabels = c.vlabels.empty() ? nullptr : c.vlabels.data();
  if (num_vertex_classes) *num_vertex_classes = hv ? c.num_vertex_classes : 0;
  if (elabels) *elabels = c.elabels.empty() ? nullptr : c.elabels.data();
  if (num_edge_classes) *num_edge_classes = he ? c.num_edge_classes : 0;
  return MGPM_OK;
}

int mgpm_graph_set_labels(mgpm_graph *g, int32_t num_vertex_classes,
                          const uint8_t *vlabels, int32_t num_edge_classes,
                          const uint16_t *elabels) {
  if (!g) return fail(MGPM_ERR_INVALID, "null argument");
  if ((!vlabels && num_vertex_classes) || (!elabels && num_edge_classes))
    return fail(MGPM_ERR_INVALID, "class count without labels");
  MGPM_GUARD_BEGIN
  CsrGraph &c = *g->g;
  CsrGraph t;  // checked before g changes
  t.n = c.n;
  t.nnz = c.nnz;
  t.num_vertex_classes = num_vertex_classes;
  t.num_edge_classes = num_edge_classes;
  if (vlabels) {
    if (!(num_vertex_classes > 0 && num_vertex_classes < 255))
      return fail(MGPM_ERR_FORMAT,
                  "vertex labels need 0 < num_vertex_classes < 255");
    t.vlabels.assign(vlabels, vlabels + c.n);
  }
  if (elabels) {
    if (!(num_edge_classes > 0 && num_edge_classes < 65535))
      return fail(MGPM_ERR_FORMAT,
                  "edge labels need 0 < num_edge_classes < 65535");
    t.elabels.assign(elabels, elabels + c.nnz);
  }
  std::string err;
  int rc = check_labels(t, err);
  if (rc) return fail(rc, err);
  c.num_vertex_classes = t.num_vertex_classes;
  c.num_edge_classes = t.num_edge_classes;
  c.vlabels.swap(t.vlabels);
  c.elabels.swap(t.elabels);
  return MGPM_OK;
  MGPM_GUARD_END
}

int mgpm_graph_label_frequency(const mgpm_graph *g, int64_t *freq) {
  if (!g || !freq) return fail(MGPM_ERR_INVALID, "null argument");
  const CsrGraph &c = *g->g;
  if (c.num_vertex_classes <= 0)
    return fail(MGPM_ERR_INVALID, "graph has no vertex labels");
  MGPM_GUARD_BEGIN
  std::vector<int64_t> f;
  if (!label_frequency(c, f))
    return fail(MGPM_ERR_FORMAT, "vertex label above num_vertex_classes");
  std::copy(f.begin(), f.end(), freq);
  return MGPM_OK;
  MGPM_GUARD_END
}

int mgpm_graph_nlf(const mgpm_graph *g, int64_t *offsets, int32_t *labels,
                   int32_t *counts) {
  if (!g || !offsets) return fail(MGPM_ERR_INVALID, "null argument");
  const CsrGraph &c = *g->g;
  if (c.num_vertex_classes <= 0)
    return fail(MGPM_ERR_INVALID, "graph has no vertex labels");
  if (!labels != !counts)
    return fail(MGPM_ERR_INVALID, "labels and counts go together");
  MGPM_GUARD_BEGIN
  nlf_csr(c, offsets, nullptr, nullptr);
  if (labels) nlf_csr(c, offsets, labels, counts);
  return MGPM_OK;
  MGPM_GUARD_END
}

int mgpm_graph_read_edges(const char *path, int32_t format, int32_t base,
                          mgpm_graph **out) {
  if (!path || !out) return fail(MGPM_ERR_INVALID, "null argument");
  if (format < 0 || format > 2 || base < -1)
    return fail(MGPM_ERR_INVALID, "bad format or base");
  MGPM_GUARD_BEGIN
  std::string p(path);
  if (format == 0) {
    const bool mtx = p.size() >= 4 && p.compare(p.size() - 4, 4, ".mtx") == 0;
    format = mtx ? 1 : 2;
  }
  if (base < 0) base = 1;
  auto g = std::make_shared<CsrGraph>();
  std::string err;
  int rc = read_edge_file(p, format, base, *g, err);
  if (rc) return fail(rc, err);
  return wrap_graph(std::move(g), out);
  MGPM_GUARD_END
}

int mgpm_graph_from_csr(int32_t n, int64_t nnz, const int64_t *rowptr,
                        const int32_t *colidx, mgpm_graph **out) {
  if (!out || n < 0 || nnz < 0 || !rowptr || (nnz > 0 && !colidx))
    return fail(MGPM_ERR_INVALID, "bad CSR arguments");
  MGPM_GUARD_BEGIN
  auto g = std::make_shared<CsrGraph>();
  g->n = n;
  g->nnz = nnz;
  g->rowptr.assign(rowptr, rowptr + n + 1);
  g->colidx.assign(colidx, colidx + nnz);
  std::string err;
  int rc = check_simple(*g, err);
  if (rc) return fail(rc, err);
  return wrap_graph(std::move(g), out);
  MGPM_GUARD_END
}

int mgpm_graph_write(const mgpm_graph *g, const char *prefix) {
  if (!g || !prefix) return fail(MGPM_ERR_INVALID, "null argument");
  MGPM_GUARD_BEGIN
  std::string err;
  int rc = write_graph(prefix, *g->g, err);
  return rc ? fail(rc, err) : MGPM_OK;
  MGPM_GUARD_END
}

int mgpm_graph_info_get(const mgpm_graph *g, mgpm_graph_info *info) {
  if (!g || !info) return fail(MGPM_ERR_INVALID, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->n = g->g->n;
  info->nnz = g->g->nnz;
  info->max_degree = g->g->max_degree();
  info->meta_max_degree = g->g->meta_max_degree;
  return MGPM_OK;
}

int mgpm_graph_csr(const mgpm_graph *g, const int64_t **rowptr,
                   const int32_t **colidx) {
  if (!g || !rowptr || !colidx) return fail(MGPM_ERR_INVALID, "null argument");
  *rowptr = g->g->rowptr.data();
  *colidx = g->g->colidx.data();
  return MGPM_OK;
}

int mgpm_graph_free(mgpm_graph *g) {
  delete g;
  return MGPM_OK;
}

int mgpm_graph_generate_rmat(int32_t scale, int32_t edge_factor, double a,
                             double b, double c, uint64_t seed,
                             mgpm_graph **out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 ||
      (int64_t(edge_factor) << scale) > (int64_t(1) << 34) || a < 0 ||
      b < 0 || c < 0 || a + b + c > 1.0)
    return fail(MGPM_ERR_INVALID, "bad RMAT parameters");
  MGPM_GUARD_BEGIN
  auto g = std::make_shared<CsrGraph>();
  generate_rmat(scale, edge_factor, a, b, c, seed, *g);
  return wrap_graph(std::move(g), out);
  MGPM_GUARD_END
}

int mgpm_graph_generate_er(int32_t n, int64_t draws, uint64_t seed,
                           mgpm_graph **out) {
  if (!out || n < 1 || draws < 0 || draws > (int64_t(1) << 34))
    return fail(MGPM_ERR_INVALID, "bad ER parameters");
  MGPM_GUARD_BEGIN
  auto g = std::make_shared<CsrGraph>();
  generate_er(n, draws, seed, *g);
  return wrap_graph(std::move(g), out);
  MGPM_GUARD_END
}

int mgpm_graph_orientation(const mgpm_graph *g, int32_t deg_thd,
                           int64_t sizes[3], int64_t *dag_rowptr,
                           int32_t *dag_colidx, int64_t *small_rowptr,
                           int32_t *small_colidx, int64_t *large_rowptr,
                           int32_t *large_colidx) {
  if (!g || !sizes) return fail(MGPM_ERR_INVALID, "null argument");
  MGPM_GUARD_BEGIN
  const CsrGraph &G = *g->g;
  Dag d;
  build_dag(G, d);
  Division dv;
  build_division(G, d, deg_thd, dv);
  sizes[0] = d.rowptr[G.n];
  sizes[1] = dv.small_rowptr[G.n];
  sizes[2] = dv.large_rowptr[G.n];
  const size_t rp = (size_t(G.n) + 1) * sizeof(int64_t);
  if (dag_rowptr) std::memcpy(dag_rowptr, d.rowptr.data(), rp);
  if (dag_colidx) std::memcpy(dag_colidx, d.colidx.data(), sizes[0] * 4);
  if (small_rowptr) std::memcpy(small_rowptr, dv.small_rowptr.data(), rp);
  if (small_colidx)
    std::memcpy(small_colidx, dv.small_colidx.data(), sizes[1] * 4);
  if (large_rowptr) std::memcpy(large_rowptr, dv.large_rowptr.data(), rp);
  if (large_colidx)
    std::memcpy(large_colidx, dv.large_colidx.data(), sizes[2] * 4);
  return MGPM_OK;
  MGPM_GUARD_END
}

// ------------------------------------------------------------ devices

int mgpm_device_count(int *count) {
  if (!count) return fail(MGPM_ERR_INVALID, "null argument");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return MGPM_OK;
}

int mgpm_device_synchronize(int device) {
  int rc = set_device(device);
  if (rc) return rc;
  MGPM_HIP(hipDeviceSynchronize());
  return MGPM_OK;
}

int mgpm_dgraph_upload(const mgpm_graph *g, int device, void *stream,
                       mgpm_dgraph **out) {
  return mgpm_dgraph_upload_ex(g, device, stream, 0, out);
}

int mgpm_dgraph_upload_ex(const mgpm_graph *g, int device, void *stream,
                          int32_t flags, mgpm_dgraph **out) {
  if (!g || !out) return fail(MGPM_ERR_INVALID, "null argument");
  if (flags & ~MGPM_UPLOAD_DEVICE_PLAN)
    return fail(MGPM_ERR_INVALID, "unknown upload flags");
  MGPM_GUARD_BEGIN
  int rc = set_device(device);
  if (rc) return rc;
  static bool inited = false;
  if (!inited) {
    if (tc_device_init()) return fail(MGPM_ERR_HIP, "kernel attribute setup");
    inited = true;
  }
  std::unique_ptr<mgpm_dgraph> dg(new mgpm_dgraph);
  dg->device = device;
  dg->host = g->g;
  MGPM_HIP(hipDeviceGetAttribute(&dg->cus, hipDeviceAttributeMultiprocessorCount,
                                 device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  MGPM_HIP(dg->scratch.alloc(256));
  MGPM_HIP(hipMemsetAsync(dg->scratch.p, 0, 256, s));
  MGPM_HIP(hipHostMalloc(reinterpret_cast<void **>(&dg->pinned), 64,
                         hipHostMallocDefault));
  MGPM_HIP(hipEventCreate(&dg->ev0));
  MGPM_HIP(hipEventCreate(&dg->ev1));
  const CsrGraph &G = *g->g;
  rc = upload_vec(dg->csr.rowptr, G.rowptr, s);
  if (rc) return rc;
  rc = upload_vec(dg->csr.colidx, G.colidx, s);
  if (rc) return rc;
  rc = (flags & MGPM_UPLOAD_DEVICE_PLAN)
           ? tc_make_resident_device(dg.get(), tc_params_from(nullptr, G.nnz / 2), s)
           : tc_make_resident(dg.get(), tc_params_from(nullptr, G.nnz / 2), s);
  if (rc) return rc;
  *out = dg.release();
  return MGPM_OK;
  MGPM_GUARD_END
}

int mgpm_dgraph_footprint(const mgpm_dgraph *dg, uint64_t *bytes) {
  if (!dg || !bytes) return fail(MGPM_ERR_INVALID, "null argument");
  *bytes = dg->tc.bytes() + dg->csr.bytes() + dg->pat.bytes() +
           dg->dag.bytes() + dg->census.bytes() + dg->scratch.bytes;
  return MGPM_OK;
}

int mgpm_dgraph_free(mgpm_dgraph *dg) {
  if (!dg) return MGPM_OK;
  (void)hipSetDevice(dg->device);
  // Every call on the handle synchronises its own stream before returning,
  // so no work of it is in flight here; batches and servers that hold the
  // graph must be done with it (their own free/destroy waits for them).
  // No device-wide synchronisation: a server freeing one graph would stall
  // every batch in flight on the device.
  if (dg->ev0) (void)hipEventDestroy(dg->ev0);
  if (dg->ev1) (void)hipEventDestroy(dg->ev1);
  if (dg->pinned) (void)hipHostFree(dg->pinned);
  delete dg;
  return MGPM_OK;
}

// ------------------------------------------------------------ counting

int mgpm_tc_count(mgpm_dgraph *dg, const mgpm_tune *tune, void *stream,
                  uint64_t *count, mgpm_stats *stats) {
  if (!dg || !count) return fail(MGPM_ERR_INVALID, "null argument");
  MGPM_GUARD_BEGIN
  const auto t0 = std::chrono::steady_clock::now();
  int rc = set_device(dg->device);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double prep = 0;
  // NULL keeps the resident plan (the defaults, or what mgpm_autotune
  // chose); an explicit tune rebuilds it when its plan knobs differ.
  const TcPlanParams want =
      tune ? tc_params_from(tune, dg->host->nnz / 2) : dg->tc.params;
  if (!same_params(want, dg->tc.params)) {
    const auto tp = std::chrono::steady_clock::now();
    rc = tc_replan(dg, want, s);
    if (rc) return rc;
    prep = ms_since(tp);
  }
  TcLaunch l;
  l.cus = dg->cus;
  l.grid_cap = (tune && tune->grid_blocks > 0) ? tune->grid_blocks
                                               : dg->tc.grid_cap;
  l.items = (tune && (tune->tc_items == 4 || tune->tc_items == 8))
                ? tune->tc_items
                : dg->tc.items;
  l.cursor_min = (tune && tune->tc_cursor_min != 0) ? tune->tc_cursor_min
                                                    : dg->tc.cursor_min;
  l.merge_max = (tune && tune->tc_merge_max != 0) ? tune->tc_merge_max
                                                  : dg->tc.merge_max;
  if (l.merge_max < 0) l.merge_max = 0;
  const int hw = (tune && tune->tc_half_waves != 0) ? tune->tc_half_waves
                                                    : dg->tc.half_waves;
  // Default: two small tasks per wave when the plan has only small tasks
  // (ER-like graphs, where per-task costs dominate).
  const TcDevicePlan &dp = dg->tc.dev;
  l.half_small = hw > 0 || (hw == 0 && dp.ngroups[kTcMedium] == 0 &&
                            dp.ngroups[kTcBlock] == 0 &&
                            dp.ngroups[kTcGlobal] == 0);
  // Laid-out partner lists: asked for (1), refused (-1), or by default
  // when they fit a quarter of the free device memory.
  const int lay = (tune && tune->tc_layout != 0) ? tune->tc_layout
                                                 : dg->tc.layout;
  bool use_layout = lay > 0;
  if (lay == 0) {
    size_t free_b = 0, total_b = 0;
    if (dg->tc.laid.p) use_layout = true;
    else if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      use_layout = tc_layout_bytes(dg->tc) <= free_b / 4;
  }
  TcDevicePlan plan = dg->tc.dev;
  if (use_layout) {
    const auto tp = std::chrono::steady_clock::now();
    const bool had = dg->tc.laid.p != nullptr;
    rc = tc_layout_build(dg, s);
    if (rc == MGPM_ERR_NOMEM && lay == 0) {
      // By default a layout that cannot be allocated is skipped: the
      // lists are gathered instead (the failed allocation's error is
      // cleared so that the launch check does not see it).
      (void)hipGetLastError();
      use_layout = false;
    } else if (rc) {
      return rc;
    } else {
      if (!had) prep += ms_since(tp);
      plan.items = dg->tc.laid.as<int32_t>();
      plan.item_off = dg->tc.laid_off.as<uint64_t>();
    }
  }
  MGPM_HIP(hipEventRecord(dg->ev0, s));
  if (tc_device_launch(plan, l, s))
    return fail(MGPM_ERR_HIP, "triangle kernel launch failed");
  MGPM_HIP(hipEventRecord(dg->ev1, s));
  MGPM_HIP(hipMemcpyAsync(dg->pinned, dg->tc.dev.result, 8,
                          hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  float kms = 0;
  MGPM_HIP(hipEventElapsedTime(&kms, dg->ev0, dg->ev1));
  *count = dg->pinned[0];
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->kernel_ms = kms;
    stats->prep_ms = prep;
    stats->n = static_cast<uint64_t>(dg->host->n);
    stats->m = static_cast<uint64_t>(dg->host->nnz / 2);
    stats->m_plus = dg->tc.m_plus;
    stats->s_plus = dg->tc.s_plus;
    stats->probes = dg->tc.probes;
    stats->bytes_alg =
        8ull * (stats->n + 1) + 20ull * stats->m_plus + 4ull * stats->s_plus;
    stats->path = use_layout ? 2 : 1;
    stats->e2e_ms = ms_since(t0);
  }
  return MGPM_OK;
  MGPM_GUARD_END
}

}  // extern "C"

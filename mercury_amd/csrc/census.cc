// P1..P8 in closed form from a subgraph census on the device.
//
// The reference counts every pattern per directed edge with set algebra
// (pattern_kernels.cuh:6-963, combined as sm_multigpu.cu:209-232).  Summed
// over the edges, each pattern's counters are a fixed combination of
// induced-subgraph counts:
//   P1 = claws K1,3, P2 = forks, P3 = 4 K1,4 + 3 claws, P4 = paths P4,
//   P5 = 2 bulls, P6 = 2 (diamond + pendant at a degree-3 vertex) + 2 paws,
//   P7 = 2 K1,1,3, P8 = 3 (K4 + pendant).
// Induced counts follow from subgraph counts N(H) (copies of H as a
// subgraph, not necessarily induced) by inverting the containment matrix
// of the graphs on 4 and 5 vertices.  Every N(H) needed is a sum the device
// forms from per-edge triangle counts t_e, per-triangle 4-clique counts k_T,
// 4-cycles and 5-cliques, and -- for wheels, K2,3 and K2,3 plus an edge --
// from the pair passes of census5_kernels.hip.
//
// P1 = sum_v C(d_v,3) - sum_triangles (d_x + d_y + d_z - 6)
//      + 2 sum_e C(t_e,2) - 4 K4                      (claws by centre)
// P4 = sum_e (d_x - t_e - 1)(d_y - t_e - 1)
//      - 4 (C4 - sum_e C(t_e,2) + 3 K4)                 (per task edge)
// P7 = 2 sum_e C(t_e,3) - sum_T k_T (t_ab + t_ac + t_bc - 6)
//      + 6 sum_T C(k_T,2) - 20 K5
// P8 = sum_T k_T (d_a + d_b + d_c - 9) - 3 sum_T k_T (t_ab + t_ac + t_bc - 6)
//      + 18 sum_T C(k_T,2) - 60 K5
// and, with the subgraph counts of census_combine below,
//   K1,4 = star - cricket + dp3 + bowtie - 2 book3 - k4pend - gem
//          + 2 k4wing + W4 - 3 k5e + 5 K5
//   fork = fork - 2 cricket - 2 bull - lolli - 2 banner + 5 dp3 + 4 bowtie
//          + 4 dp2 + 4 house + 6 K23 - 12 book3 - 9 k4pend - 10 gem
//          - 10 K23e + 20 k4wing + 20 W4 - 36 k5e + 60 K5
//   bull = bull - 2 dp3 - 2 dp2 - house + 6 book3 + 6 k4pend + 5 gem
//          + 4 K23e - 14 k4wing - 12 W4 + 30 k5e - 60 K5
//   dp3i = dp3 - 6 book3 - 3 k4pend - 2 gem + 10 k4wing + 4 W4
//          - 24 k5e + 60 K5
//   paw  = sum_T (d_a + d_b + d_c - 6) - 4 sum_e C(t_e,2) + 12 K4.
// The identities are checked term by term against brute-force induced
// counts in tests/test_census5_identity.py (and census_identity.py).
// Counts are uint64 modulo 2^64 throughout, like the reference's unsigned
// counters; no intermediate sum is divided except by 3, through its
// inverse modulo 2^64 (exact for a multiple of 3 whatever the wrap).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "capi_common.hpp"
#include "census.hpp"
#include "dag_device.hpp"

namespace mgpm {

namespace {

constexpr int kMaxCensusDevices = 64;

// Per-caller budgets of the per-workgroup arrays.
constexpr size_t kCensusC4Budget = size_t(4) << 30;
constexpr size_t kCensusK5Budget = size_t(8) << 30;
constexpr size_t kCensus5Budget = size_t(8) << 30;  // pair-table layout

// Layout of the small buffer: census cursors, accumulators, pair-pass
// cursors and class counters.
constexpr size_t kSmallBytes = 1024;
constexpr size_t kAccOff = 64;
constexpr size_t kC5CursorOff = 384;
constexpr size_t kClsCountOff = 448;
constexpr size_t kClsFillOff = 480;

// `passes` is a census_passes mask: the triangle pass and the sums always
// run, the 4-cycle, 5-clique and pair passes on request; the pair passes
// also add the 5-vertex terms to the others.
bool is_k5(int passes) { return (passes & kCensusPassK5) != 0; }
bool is_full5(int passes) { return (passes & kCensusPassPairs) != 0; }
bool needs_c4(int passes) { return (passes & kCensusPassC4) != 0; }

// The inverse of 3 modulo 2^64.
constexpr uint64_t kInv3 = 0xAAAAAAAAAAAAAAABull;

template <class T>
int grow(DevBuf &b, size_t count) {
  const size_t want = std::max<size_t>(count, 1) * sizeof(T);
  if (b.bytes >= want) return MGPM_OK;
  MGPM_HIP(b.alloc(want));
  return MGPM_OK;
}

int zero(void *p, size_t bytes, hipStream_t s) {
  if (!p || !bytes) return MGPM_OK;
  if (device_zero(p, (bytes + 3) / 4, s))
    return fail(MGPM_ERR_HIP, "census fill kernel launch failed");
  return MGPM_OK;
}

// Sizes of the workspace for one pattern on one graph.
struct Need {
  size_t t = 0, t_csr = 0, tv = 0, slab = 0, wedges = 0, k5 = 0;
  size_t stamp = 0, mask = 0, grows = 0, anchors = 0, table = 0;
  int c4_groups = 0, k5_waves = 0, groups = 0;
  int64_t k5_len = 0, mask_words = 0, grows_len = 0;
};

Need need_for(const mgpm_dgraph *dg, int passes, size_t budget) {
  const CsrGraph &G = *dg->host;
  const int64_t n = std::max<int32_t>(G.n, 1);
  const int64_t m = std::max<int64_t>(G.nnz / 2, 1);
  const int64_t nnz = std::max<int64_t>(G.nnz, 1);
  const CensusGraph &cg = dg->census;
  Need k;
  k.t = static_cast<size_t>(m);
  k.slab = static_cast<size_t>(census_waves(dg->cus)) * cg.max_out;
  if (needs_c4(passes)) {
    const size_t per = static_cast<size_t>(n) * sizeof(uint32_t);
    k.c4_groups = static_cast<int>(std::min<size_t>(
        1024, std::max<size_t>(64, kCensusC4Budget / (2 * per))));
    k.wedges = static_cast<size_t>(k.c4_groups) * n;
  }
  if (is_k5(passes)) {
    k.k5_len = census_k5_slab_len(cg.max_out, cg.max_deg);
    const size_t per = static_cast<size_t>(k.k5_len) * sizeof(int32_t);
    const size_t fit = kCensusK5Budget / per;
    k.k5_waves = static_cast<int>(
        std::max<size_t>(4, std::min<size_t>(census_waves(dg->cus), fit)) &
        ~size_t(3));
    k.k5 = static_cast<size_t>(k.k5_waves) * k.k5_len;
  }
  if (is_full5(passes)) {
    k.t_csr = static_cast<size_t>(nnz);
    k.tv = static_cast<size_t>(n);
    const int gen_words = static_cast<int>((cg.max_out + 63) / 64);
    // Half the budget to stamps, pending counts and lists (12 bytes per
    // vertex per workgroup), half to masks: the one- and two-word classes
    // (nearly every anchor) get every workgroup, wider classes fewer.
    const size_t half = budget / 2;
    k.groups = static_cast<int>(std::max<size_t>(
        1, std::min<size_t>({size_t(dg->cus) * 4, size_t(n),
                             half / (12 * size_t(n))})));
    const int64_t wmax = std::max<int64_t>(8, gen_words);
    k.mask_words = std::min<int64_t>(half / (8 * size_t(n)), 2 * k.groups);
    k.mask_words = std::max<int64_t>(k.mask_words, wmax);
    k.stamp = static_cast<size_t>(k.groups) * n;
    k.mask = static_cast<size_t>(k.mask_words) * n;
    k.grows_len = cg.max_out * gen_words;
    k.grows = static_cast<size_t>(k.groups) * k.grows_len;
    k.anchors = static_cast<size_t>(n);
    k.table = cg.tab_total;
  }
  return k;
}

// Builds the census graph on the device, once: the graph relabelled in DAG
// order, then its DAG.
int graph_ready(mgpm_dgraph *dg, hipStream_t s) {
  CensusGraph &cg = dg->census;
  if (cg.ready) return MGPM_OK;
  const CsrGraph &G = *dg->host;
  const int64_t nnz = G.nnz;
  const int32_t n = G.n;
  const int64_t m = nnz / 2;
  if (nnz >= (int64_t(1) << 32))
    return fail(MGPM_ERR_CAPACITY, "census: 2^32 or more arcs");
  CensusGraph g;
  MGPM_HIP(g.rowptr.alloc((static_cast<size_t>(n) + 1) * sizeof(int64_t)));
  MGPM_HIP(g.col.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  {
    const size_t rb = census_relabel_bytes(n, nnz);
    if (rb == 0) return fail(MGPM_ERR_HIP, "sort size query");
    DevBuf rtmp;
    MGPM_HIP(rtmp.alloc(rb));
    if (census_relabel(dg->csr.rowptr.as<int64_t>(),
                       dg->csr.colidx.as<int32_t>(), n, nnz,
                       g.rowptr.as<int64_t>(), g.col.as<int32_t>(), rtmp.p,
                       rb, dg->cus, s))
      return fail(MGPM_ERR_HIP, "census relabelling failed");
    MGPM_HIP(hipStreamSynchronize(s));  // rtmp is freed here
  }
  DevBuf nd, val, pos, scan;
  const size_t a1 = static_cast<size_t>(nnz) + 1;
  MGPM_HIP(g.deg.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(nd.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(g.src.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  MGPM_HIP(val.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(pos.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(g.drp.alloc((static_cast<size_t>(n) + 1) * sizeof(int64_t)));
  MGPM_HIP(g.dcol.alloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  MGPM_HIP(g.dsrc.alloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  size_t tb = 0;
  if (dag_scan_bytes(nnz, &tb)) return fail(MGPM_ERR_HIP, "scan size query");
  MGPM_HIP(scan.alloc(std::max<size_t>(tb, 1)));
  DagDevice d;
  d.rowptr = g.rowptr.as<int64_t>();
  d.col = g.col.as<int32_t>();
  d.n = n;
  d.nnz = nnz;
  d.division = false;
  d.cus = dg->cus;
  d.deg = g.deg.as<int32_t>();
  d.nd = nd.as<int32_t>();
  d.src = g.src.as<int32_t>();
  d.val = val.as<int64_t>();
  d.pos = pos.as<int64_t>();
  d.scan_tmp = scan.p;
  d.scan_bytes = tb;
  d.out_rowptr[0] = g.drp.as<int64_t>();
  d.out_col[0] = g.dcol.as<int32_t>();
  if (dag_device_build(d, s)) return fail(MGPM_ERR_HIP, "DAG kernel launch failed");
  if (census_dag_sources(g.drp.as<int64_t>(), n, g.dsrc.as<int32_t>(),
                         dg->cus, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  g.max_deg = std::max<int32_t>(G.max_degree(), 1);
  // The host plan knows the largest DAG out-degree (the same DAG).
  g.max_out = std::max<int64_t>(dg->tc.max_out, 1);
  MGPM_HIP(hipStreamSynchronize(s));  // the work arrays die here
  cg.rowptr.swap(g.rowptr);
  cg.col.swap(g.col);
  cg.deg.swap(g.deg);
  cg.src.swap(g.src);
  cg.drp.swap(g.drp);
  cg.dcol.swap(g.dcol);
  cg.dsrc.swap(g.dsrc);
  cg.max_deg = g.max_deg;
  cg.max_out = g.max_out;
  cg.ready = true;
  return MGPM_OK;
}

unsigned *cursor_of(const CensusWork &w) { return w.small.as<unsigned>(); }

CensusArgs args_of(mgpm_dgraph *dg, int passes, CensusWork &w,
                   const Need &k) {
  const CensusGraph &cg = dg->census;
  CensusArgs c;
  c.drp = cg.drp.as<int64_t>();
  c.dcol = cg.dcol.as<int32_t>();
  c.dsrc = cg.dsrc.as<int32_t>();
  c.deg = cg.deg.as<int32_t>();
  c.n = dg->host->n;
  c.m = dg->host->nnz / 2;
  c.t = w.t.as<uint32_t>();
  c.slab = w.slab.as<int32_t>();
  c.slab_len = cg.max_out;
  c.cursor = cursor_of(w);
  c.acc = census_acc(w);
  c.rowptr = cg.rowptr.as<int64_t>();
  c.col = cg.col.as<int32_t>();
  c.wedges = w.wedges.as<uint32_t>();
  c.touched = w.touched.as<int32_t>();
  c.c4_groups = k.c4_groups;
  c.k5_slab = w.k5_slab.as<int32_t>();
  c.k5_len = k.k5_len;
  c.max_deg = cg.max_deg;
  c.k5_waves = k.k5_waves;
  c.full5 = is_full5(passes);
  c.t_csr = w.t_csr.as<uint32_t>();
  c.tv = w.tv.as<unsigned long long>();
  return c;
}

Census5Args args5_of(mgpm_dgraph *dg, CensusWork &w, const Need &k) {
  const CensusGraph &cg = dg->census;
  Census5Args c;
  c.rowptr = cg.rowptr.as<int64_t>();
  c.col = cg.col.as<int32_t>();
  c.src = cg.src.as<int32_t>();
  c.deg = cg.deg.as<int32_t>();
  c.n = dg->host->n;
  c.nnz = dg->host->nnz;
  c.drp = cg.drp.as<int64_t>();
  c.dcol = cg.dcol.as<int32_t>();
  c.t = w.t.as<uint32_t>();
  c.t_csr = w.t_csr.as<uint32_t>();
  c.tv = w.tv.as<unsigned long long>();
  c.anchors = w.anchors.as<int32_t>();
  char *sm = w.small.as<char>();
  c.cls_count = reinterpret_cast<unsigned *>(sm + kClsCountOff);
  c.cls_fill = reinterpret_cast<unsigned *>(sm + kClsFillOff);
  c.stamp = w.stamp.as<uint32_t>();
  c.lo = w.lo.as<uint32_t>();
  c.mask = w.mask.as<unsigned long long>();
  c.list = w.list.as<int32_t>();
  c.groups = k.groups;
  c.mask_words = k.mask_words;
  c.gen_words = static_cast<int>((cg.max_out + 63) / 64);
  c.grows = w.grows.as<unsigned long long>();
  c.grows_len = k.grows_len;
  c.tab_off = cg.tab_off.as<unsigned long long>();
  c.table = w.table.as<unsigned long long>();
  c.cursor = reinterpret_cast<unsigned *>(sm + kC5CursorOff);
  c.acc = census_acc(w);
  return c;
}

// Grows `w` to hold `k` (every buffer grow-only by bytes, so each query's
// own layout -- derived from its graph by need_for at enqueue -- fits
// whatever larger queries grew it to).  Arrays the kernels leave zeroed
// are zeroed when (re)allocated; the rest is filled per call.
int work_grow(CensusWork &w, const Need &k, hipStream_t s) {
  int rc;
  auto grow_z = [&](DevBuf &b, size_t bytes) -> int {
    if (b.bytes >= std::max<size_t>(bytes, 1)) return MGPM_OK;
    MGPM_HIP(b.alloc(std::max<size_t>(bytes, 1)));
    return zero(b.p, b.bytes, s);
  };
  if ((rc = grow<uint32_t>(w.t, k.t))) return rc;
  if ((rc = grow<char>(w.small, kSmallBytes))) return rc;
  if ((rc = grow<int32_t>(w.slab, k.slab))) return rc;
  if (k.wedges) {
    if ((rc = grow_z(w.wedges, k.wedges * 4))) return rc;
    if ((rc = grow<int32_t>(w.touched, k.wedges))) return rc;
  }
  if (k.k5 && (rc = grow<int32_t>(w.k5_slab, k.k5))) return rc;
  if (k.t_csr) {
    if ((rc = grow<uint32_t>(w.t_csr, k.t_csr))) return rc;
    if ((rc = grow<unsigned long long>(w.tv, k.tv))) return rc;
    if ((rc = grow<int32_t>(w.anchors, k.anchors))) return rc;
    if ((rc = grow<unsigned long long>(w.table, k.table))) return rc;
    if ((rc = grow<uint32_t>(w.stamp, k.stamp))) return rc;
    if ((rc = grow<int32_t>(w.list, k.stamp))) return rc;
    if ((rc = grow_z(w.lo, k.stamp * 4))) return rc;
    if ((rc = grow_z(w.mask, k.mask * 8))) return rc;
    if ((rc = grow<unsigned long long>(w.grows, k.grows))) return rc;
  }
  return MGPM_OK;
}

}  // namespace

unsigned long long *census_acc(const CensusWork &w) {
  return reinterpret_cast<unsigned long long *>(w.small.as<char>() + kAccOff);
}

uint64_t census_work_bytes(const mgpm_dgraph *dg, int passes,
                           size_t budget) {
  if (!dg->census.ready) {
    // The DAG the graph will keep, plus a bound on the workspace.
    const CsrGraph &G = *dg->host;
    const uint64_t dag = 4ull * G.n + 4ull * G.nnz + 8ull * (G.n + 1) +
                         8ull * (G.nnz / 2);
    const uint64_t work = is_full5(passes) ? budget + 16ull * G.nnz
                                           : 32ull * G.nnz;
    return dag + work;
  }
  const Need k = need_for(dg, passes, budget);
  return 4ull * k.t + kSmallBytes + 4ull * k.slab + 8ull * k.wedges +
         4ull * k.k5 + 4ull * k.t_csr + 8ull * k.tv + 12ull * k.stamp +
         8ull * k.mask + 8ull * k.grows + 4ull * k.anchors + 8ull * k.table;
}

namespace {

// The pair-table layout of the graph (once): per anchor, a power of two of
// slots at least twice the pairs it will hold, from one count pass over
// temporary stamps, one scan and one read of the total.
int pairs_layout(mgpm_dgraph *dg, hipStream_t s) {
  CensusGraph &cg = dg->census;
  if (cg.pairs_ready) return MGPM_OK;
  const int32_t n = dg->host->n;
  const int64_t nnz = dg->host->nnz;
  const size_t n1 = static_cast<size_t>(n) + 1;
  const int groups = static_cast<int>(std::max<size_t>(
      1, std::min<size_t>(size_t(dg->cus) * 2,
                          (kCensus5Budget / 4) /
                              (4 * std::max<size_t>(n, 1)))));
  DevBuf off, cap, stamp, small, scan;
  size_t sb = 0;
  if (census5_scan_bytes(int64_t(n1), &sb))
    return fail(MGPM_ERR_HIP, "scan size query");
  MGPM_HIP(off.alloc(n1 * 8));
  MGPM_HIP(cap.alloc(n1 * 8));
  MGPM_HIP(stamp.alloc(size_t(groups) * std::max<size_t>(n, 1) * 4));
  MGPM_HIP(small.alloc(kSmallBytes));
  MGPM_HIP(scan.alloc(std::max<size_t>(sb, 1)));
  Census5Args c;
  c.rowptr = cg.rowptr.as<int64_t>();
  c.col = cg.col.as<int32_t>();
  c.deg = cg.deg.as<int32_t>();
  c.n = n;
  c.nnz = nnz;
  c.drp = cg.drp.as<int64_t>();
  c.dcol = cg.dcol.as<int32_t>();
  c.stamp = stamp.as<uint32_t>();
  c.groups = groups;
  c.tab_cap = cap.as<unsigned long long>();
  c.cursor = reinterpret_cast<unsigned *>(small.as<char>() + kC5CursorOff);
  int rc;
  if ((rc = zero(small.p, kSmallBytes, s))) return rc;
  if ((rc = zero(stamp.p, stamp.bytes, s))) return rc;
  if ((rc = zero(cap.p, n1 * 8, s))) return rc;
  if (census5_table_count_launch(c, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  if (census5_scan(cap.as<unsigned long long>(), off.as<unsigned long long>(),
                   int64_t(n1), scan.p, scan.bytes, s))
    return fail(MGPM_ERR_HIP, "census scan failed");
  unsigned long long total = 0;
  MGPM_HIP(hipMemcpyAsync(&total, off.as<unsigned long long>() + n, 8,
                          hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  cg.tab_off.swap(off);
  cg.tab_total = total;
  cg.pairs_ready = true;
  return MGPM_OK;
}

}  // namespace

int census_graph_prepare(mgpm_dgraph *dg, hipStream_t s) {
  int rc = graph_ready(dg, s);
  if (rc) return rc;
  return pairs_layout(dg, s);
}

int census_passes(int pattern) {
  switch (pattern) {
    case MGPM_P1: return 0;
    case MGPM_P4: return kCensusPassC4;
    case MGPM_P7: case MGPM_P8: return kCensusPassK5;
    default:  // P2, P3, P5, P6
      return kCensusPassC4 | kCensusPassK5 | kCensusPassPairs;
  }
}

int census_prepare(mgpm_dgraph *dg, int passes, CensusWork &w,
                   hipStream_t s) {
  int rc = graph_ready(dg, s);
  if (rc) return rc;
  if (is_full5(passes) && (rc = pairs_layout(dg, s))) return rc;
  return work_grow(w, need_for(dg, passes, w.budget), s);
}

int census_enqueue(mgpm_dgraph *dg, int passes, CensusWork &w,
                   hipStream_t s) {
  const Need k = need_for(dg, passes, w.budget);
  CensusArgs c = args_of(dg, passes, w, k);
  int rc;
  if ((rc = zero(w.t.p, static_cast<size_t>(c.m) * 4, s))) return rc;
  if ((rc = zero(w.small.p, kSmallBytes, s))) return rc;
  const bool f5 = is_full5(passes);
  if (f5 && (rc = zero(w.tv.p, static_cast<size_t>(c.n) * 8, s))) return rc;
  if (census_launch(c, dg->cus, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  Census5Args c5;
  if (f5) {
    c5 = args5_of(dg, w, k);
    if (census5_flat_launch(c5, dg->cus, s))
      return fail(MGPM_ERR_HIP, "census kernel launch failed");
  }
  if (needs_c4(passes) && census_c4_launch(c, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  if (is_k5(passes) && census_k5_launch(c, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  if (f5) {
    const CensusGraph &cg = dg->census;
    if ((rc = zero(w.table.p, cg.tab_total * 8, s))) return rc;
    if (census5_classify_launch(c5, s))
      return fail(MGPM_ERR_HIP, "census kernel launch failed");
    if ((rc = zero(w.stamp.p, w.stamp.bytes, s))) return rc;
    if (census5_pairs_launch(c5, s))
      return fail(MGPM_ERR_HIP, "census kernel launch failed");
  }
  return MGPM_OK;
}

int census_combine(int pattern, const unsigned long long *acc,
                   uint64_t *count) {
  const uint64_t pairs_t = acc[kCensusEdgePairs], k4 = acc[kCensusK4];
  const uint64_t tdeg = acc[kCensusTriDegrees];
  if (pattern == MGPM_P1) {
    *count = acc[kCensusDeg3] - tdeg + 2 * pairs_t - 4 * k4;
    return MGPM_OK;
  }
  if (pattern == MGPM_P4) {  // induced 4-cycles = C4 - sum C(t_e,2) + 3 K4
    *count = acc[kCensusPathPairs] - 4 * (acc[kCensusC4] - pairs_t + 3 * k4);
    return MGPM_OK;
  }
  // Two independent counts of the 4-cliques, every common neighbourhood of
  // the size its triangle count says, the halved sum even (parity survives
  // the wrap modulo 2^64), and every pair lookup found.
  if (acc[kCensusMismatch] != 0 || acc[kCensusK4Rows] != k4 ||
      acc[kCensusTriEdgeK] % 2 != 0 || acc[kCensusTableMiss] != 0)
    return fail(MGPM_ERR_INTERNAL, "census: self-checks disagree");
  const uint64_t tek = acc[kCensusTriEdgeK];  // 2 k4wing
  const uint64_t tdk = acc[kCensusTriDegK];   // 3 k4pend
  const uint64_t k5e = acc[kCensusTriPairs], k5 = acc[kCensusK5];
  const uint64_t book3 = acc[kCensusEdgeTriples];
  if (pattern == MGPM_P7) {
    *count = 2 * book3 - tek + 6 * k5e - 20 * k5;
    return MGPM_OK;
  }
  if (pattern == MGPM_P8) {
    *count = tdk - 3 * tek + 18 * k5e - 60 * k5;
    return MGPM_OK;
  }
  const uint64_t tri = acc[kCensusTriangles];
  const uint64_t k4pend = tdk * kInv3;
  const uint64_t gem = acc[kCensusGem] - 12 * k4;
  const uint64_t w4 = acc[kCensusW4];
  const uint64_t dp3 = acc[kCensusDp3];
  const uint64_t cricket = acc[kCensusCricket];
  const uint64_t bowtie = acc[kCensusTv2] - 2 * pairs_t;
  const uint64_t star = acc[kCensusDeg4] - cricket + dp3 + bowtie -
                        2 * book3 - k4pend - gem + tek + w4 - 3 * k5e +
                        5 * k5;
  const uint64_t dp3i = dp3 - 6 * book3 - 3 * k4pend - 2 * gem + 5 * tek +
                        4 * w4 - 24 * k5e + 60 * k5;
  if (pattern == MGPM_P3) {
    const uint64_t claw = acc[kCensusDeg3] - tdeg + 2 * pairs_t - 4 * k4;
    *count = 4 * star + 3 * claw;
    return MGPM_OK;
  }
  if (pattern == MGPM_P6) {
    const uint64_t paw = tdeg - 4 * pairs_t + 12 * k4;
    *count = 2 * dp3i + 2 * paw;
    return MGPM_OK;
  }
  const uint64_t bull = acc[kCensusBullDeg] - 2 * pairs_t;
  const uint64_t dp2 = acc[kCensusDp2] - 12 * k4;
  const uint64_t house = acc[kCensusCycEdgeT] - 4 * pairs_t;
  const uint64_t k23e = acc[kCensusK23e];
  if (pattern == MGPM_P5) {
    const uint64_t bulli = bull - 2 * dp3 - 2 * dp2 - house + 6 * book3 +
                           6 * k4pend + 5 * gem + 4 * k23e - 7 * tek -
                           12 * w4 + 30 * k5e - 60 * k5;
    *count = 2 * bulli;
    return MGPM_OK;
  }
  // P2: induced forks.  sum_T (d sum) = tdeg + 6 triangles.
  const uint64_t lolli =
      acc[kCensusTriS] - 2 * (tdeg + 6 * tri) + 6 * tri - 4 * pairs_t;
  const uint64_t banner = acc[kCensusBanner] - 2 * pairs_t;
  *count = acc[kCensusFork] - 2 * cricket - 2 * bull - lolli - 2 * banner +
           5 * dp3 + 4 * bowtie + 4 * dp2 + 4 * house + 6 * acc[kCensusK23] -
           12 * book3 - 9 * k4pend - 10 * gem - 10 * k23e + 10 * tek +
           20 * w4 - 36 * k5e + 60 * k5;
  return MGPM_OK;
}

namespace {

// The single-query path's workspace: one per device, shared by every
// graph on it (a workspace per graph would hold gigabytes per handle), so
// single-query census calls on one device run one at a time.  Never freed:
// it lives until the process ends.
struct SharedWork {
  std::mutex mu;
  CensusWork w;
};

SharedWork &shared_work(int device) {
  static SharedWork *all = new SharedWork[kMaxCensusDevices];
  return all[device];
}

}  // namespace

int census_pattern_count(mgpm_dgraph *dg, int pattern, hipStream_t s,
                         uint64_t *count, float *kernel_ms) {
  if (pattern < MGPM_P1 || pattern > MGPM_P8)
    return fail(MGPM_ERR_PATTERN, "Not support.");
  if (dg->device < 0 || dg->device >= kMaxCensusDevices)
    return fail(MGPM_ERR_NODEVICE, "census: device index out of range");
  SharedWork &sw = shared_work(dg->device);
  std::lock_guard<std::mutex> lk(sw.mu);
  CensusWork &w = sw.w;
  const int passes = census_passes(pattern);
  int rc = census_prepare(dg, passes, w, s);
  if (rc) return rc;
  MGPM_HIP(hipEventRecord(dg->ev0, s));
  if ((rc = census_enqueue(dg, passes, w, s))) return rc;
  MGPM_HIP(hipEventRecord(dg->ev1, s));
  unsigned long long acc[kCensusSlots];
  MGPM_HIP(hipMemcpyAsync(acc, census_acc(w), sizeof(acc),
                          hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  MGPM_HIP(hipEventElapsedTime(kernel_ms, dg->ev0, dg->ev1));
  return census_combine(pattern, acc, count);
}

}  // namespace mgpm

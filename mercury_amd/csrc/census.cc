// P1 (induced claws), P4 (induced paths on 4 vertices), P7 (induced
// K1,1,3) and P8 (induced K4 plus a pendant edge) in closed form from a
// triangle, 4-cycle, 4-clique and 5-clique census on the device.
//
// The reference counts P1 per directed edge with frequency arrays
// (pattern_kernels.cuh:6-133, combined as sm_multigpu.cu:209-232); its sum
// is the number of induced claws K1,3.  A claw centred at v is an
// independent triple of N(v), and by inclusion-exclusion over the edges
// inside N(v):
//   claws(v) = C(d,3) - t_v (d - 2) + sum_{u in N(v)} C(t_uv, 2) - K4(v)
// (t_v: triangles at v, t_uv: triangles on edge uv, K4(v): 4-cliques at v).
// Summed over v:
//   P1 = sum_v C(d_v,3) - sum_triangles (d_x + d_y + d_z - 6)
//        + 2 sum_edges C(t_e,2) - 4 K4,
// which census_kernels.hip computes by one enumeration of the triangles of
// the degree-oriented DAG.
//
// P4 (pattern_kernels.cuh:855-963, one task per undirected edge) counts,
// per edge {x,y}, pairs a in N(x)\N[y], b in N(y)\N[x] with a, b not
// adjacent: (d_x - t - 1)(d_y - t - 1) pairs minus those with a ~ b, which
// are exactly the induced 4-cycles x-a-b-y through the edge.  Each induced
// 4-cycle has 4 edges, and 4-cycle subgraphs sit on induced 4-cycles (1),
// diamonds (1) and K4s (3), with diamonds = sum C(t_e,2) - 6 K4:
//   P4 = sum_edges (d_x - t - 1)(d_y - t - 1)
//        - 4 (C4 - sum_edges C(t_e,2) + 3 K4).
//
// P7 (pattern_kernels.cuh:352-459) is twice the induced K1,1,3: per edge
// e, the independent triples of its common neighbourhood C_e (|C_e| =
// t_e).  Inclusion-exclusion over the three possible edges of a triple
// gives C(t,3) - (t-2) E(C_e) + sum_{c in C_e} C(d_C(c), 2) - T(C_e),
// where E(C_e) = k4_e (4-cliques on e), d_C(c) = k_T (4-cliques on the
// triangle T = e + c) and T(C_e) = 5-cliques on e.  Summed over edges, and
// with sum_e (t_e - 2) k4_e = 1/2 sum_T k_T (t_ab + t_ac + t_bc - 6):
//   P7 = 2 (sum_e C(t_e,3) - 1/2 sum_T k_T (sum_{e in T} t_e - 6)
//           + 3 sum_T C(k_T,2) - 10 K5).
// P8 (pattern_kernels.cuh:461-572) is three times the induced K4 plus a
// pendant edge: per 4-clique Q and a in Q, the neighbours of a outside Q
// adjacent to no other vertex of Q, by inclusion-exclusion over those
// three: (d_a - 3) - sum_b (t_ab - 2) + sum_{b,c} (k_abc - 1) - K5(Q).
// Summed over Q and a, with sum_v (d_v - 3) k4_v = 1/3 sum_T k_T
// (d_a + d_b + d_c - 9):
//   P8 = 3 (1/3 sum_T k_T (d_a + d_b + d_c - 9)
//           - sum_T k_T (sum_{e in T} t_e - 6) + 6 sum_T C(k_T,2) - 20 K5).
// Exact in uint64 (mod 2^64 in the intermediate terms, like the
// reference's unsigned counters).
#include <algorithm>
#include <cstring>

#include "capi_common.hpp"
#include "census.hpp"
#include "dag_device.hpp"

namespace mgpm {

// Device bytes of the 4-cycle pass's per-workgroup arrays, and of the
// 5-clique pass's per-wave slabs.
constexpr size_t kCensusC4Budget = size_t(4) << 30;
constexpr size_t kCensusK5Budget = size_t(8) << 30;

int census_pattern_count(mgpm_dgraph *dg, int pattern, hipStream_t s,
                         uint64_t *count, float *kernel_ms) {
  if (pattern != MGPM_P1 && pattern != MGPM_P4 && pattern != MGPM_P7 &&
      pattern != MGPM_P8)
    return fail(MGPM_ERR_PATTERN,
                "census: only P1, P4, P7 and P8 have closed forms");
  const bool k5 = pattern == MGPM_P7 || pattern == MGPM_P8;
  const CsrGraph &G = *dg->host;
  const int64_t nnz = G.nnz;
  const int32_t n = G.n;
  const int64_t m = nnz / 2;
  if (nnz >= (int64_t(1) << 32))
    return fail(MGPM_ERR_CAPACITY, "census: 2^32 or more arcs");
  const size_t a1 = static_cast<size_t>(nnz) + 1;
  DevBuf deg, nd, src, val, pos, drp, dcol, dsrc, t, slab, small, scan;
  MGPM_HIP(deg.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(nd.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(src.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  MGPM_HIP(val.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(pos.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(drp.alloc((static_cast<size_t>(n) + 1) * sizeof(int64_t)));
  MGPM_HIP(dcol.alloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  MGPM_HIP(dsrc.alloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  MGPM_HIP(t.alloc(std::max<int64_t>(m, 1) * sizeof(uint32_t)));
  size_t tb = 0;
  if (dag_scan_bytes(nnz, &tb)) return fail(MGPM_ERR_HIP, "scan size query");
  MGPM_HIP(scan.alloc(std::max<size_t>(tb, 1)));
  // The largest DAG out-degree bounds every triangle set C.
  const int64_t slab_len = std::max<int32_t>(dg->tc.max_out, 1);
  const int waves = census_waves(dg->cus);
  MGPM_HIP(slab.alloc(static_cast<size_t>(waves) * slab_len * sizeof(int32_t)));
  MGPM_HIP(small.alloc(256));
  // P4 also counts 4-cycles: per workgroup a dense counter array and a
  // list over all vertices (within a fixed budget).
  DevBuf wedges, touched;
  int groups = 0;
  if (pattern == MGPM_P4) {
    const size_t per = std::max<size_t>(n, 1) * sizeof(uint32_t);
    groups = static_cast<int>(std::min<size_t>(
        1024, std::max<size_t>(64, kCensusC4Budget / (2 * per))));
    MGPM_HIP(wedges.alloc(groups * per));
    MGPM_HIP(touched.alloc(groups * per));
  }

  // P7 and P8 also take the 4-cliques on every triangle and the
  // 5-cliques: per wave a slab sized by the largest DAG out-degree and the
  // largest degree.
  DevBuf k5_slab;
  int32_t max_deg = 1;
  int64_t k5_len = 0;
  int k5_waves = 0;
  if (k5) {
    for (int32_t v = 0; v < n; ++v)
      max_deg = std::max<int32_t>(
          max_deg, static_cast<int32_t>(G.rowptr[v + 1] - G.rowptr[v]));
    k5_len = census_k5_slab_len(slab_len, max_deg);
    const size_t per = static_cast<size_t>(k5_len) * sizeof(int32_t);
    const size_t fit = kCensusK5Budget / per;
    k5_waves = static_cast<int>(
        std::max<size_t>(4, std::min<size_t>(waves, fit)) & ~size_t(3));
    MGPM_HIP(k5_slab.alloc(static_cast<size_t>(k5_waves) * per));
  }

  DagDevice d;
  d.rowptr = dg->csr.rowptr.as<int64_t>();
  d.col = dg->csr.colidx.as<int32_t>();
  d.n = n;
  d.nnz = nnz;
  d.division = false;
  d.cus = dg->cus;
  d.deg = deg.as<int32_t>();
  d.nd = nd.as<int32_t>();
  d.src = src.as<int32_t>();
  d.val = val.as<int64_t>();
  d.pos = pos.as<int64_t>();
  d.scan_tmp = scan.p;
  d.scan_bytes = tb;
  d.out_rowptr[0] = drp.as<int64_t>();
  d.out_col[0] = dcol.as<int32_t>();

  CensusArgs c;
  c.drp = d.out_rowptr[0];
  c.dcol = d.out_col[0];
  c.dsrc = dsrc.as<int32_t>();
  c.deg = d.deg;
  c.n = n;
  c.m = m;
  c.t = t.as<uint32_t>();
  c.slab = slab.as<int32_t>();
  c.slab_len = slab_len;
  c.cursor = small.as<unsigned>();
  c.acc = reinterpret_cast<unsigned long long *>(small.as<char>() + 64);
  c.rowptr = d.rowptr;
  c.col = d.col;
  c.wedges = wedges.as<uint32_t>();
  c.touched = touched.as<int32_t>();
  c.c4_groups = groups;
  c.k5_slab = k5_slab.as<int32_t>();
  c.k5_len = k5_len;
  c.max_deg = max_deg;
  c.k5_waves = k5_waves;

  MGPM_HIP(hipEventRecord(dg->ev0, s));
  if (dag_device_build(d, s)) return fail(MGPM_ERR_HIP, "DAG kernel launch failed");
  MGPM_HIP(hipMemsetAsync(t.p, 0, t.bytes, s));
  MGPM_HIP(hipMemsetAsync(small.p, 0, small.bytes, s));
  if (census_launch(c, dg->cus, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  if (pattern == MGPM_P4) {
    MGPM_HIP(hipMemsetAsync(wedges.p, 0, wedges.bytes, s));
    if (census_c4_launch(c, s))
      return fail(MGPM_ERR_HIP, "census kernel launch failed");
  }
  if (k5 && census_k5_launch(c, s))
    return fail(MGPM_ERR_HIP, "census kernel launch failed");
  MGPM_HIP(hipEventRecord(dg->ev1, s));
  unsigned long long acc[kCensusSlots];
  MGPM_HIP(hipMemcpyAsync(acc, c.acc, sizeof(acc), hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  MGPM_HIP(hipEventElapsedTime(kernel_ms, dg->ev0, dg->ev1));
  if (pattern == MGPM_P1) {
    *count = acc[kCensusDeg3] - acc[kCensusTriDegrees] +
             2 * acc[kCensusEdgePairs] - 4 * acc[kCensusK4];
  } else if (pattern == MGPM_P4) {  // induced 4-cycles = C4 - sum C(t_e,2) + 3 K4
    *count = acc[kCensusPathPairs] -
             4 * (acc[kCensusC4] - acc[kCensusEdgePairs] + 3 * acc[kCensusK4]);
  } else {
    // Two independent counts of the 4-cliques, every common neighbourhood
    // of the size its triangle count says, and the halving / third exact.
    if (acc[kCensusMismatch] != 0 || acc[kCensusK4Rows] != acc[kCensusK4] ||
        acc[kCensusTriEdgeK] % 2 != 0 || acc[kCensusTriDegK] % 3 != 0)
      return fail(MGPM_ERR_INTERNAL, "census: 4-clique terms disagree");
    const uint64_t tek = acc[kCensusTriEdgeK], kk = acc[kCensusTriPairs],
                   k5c = acc[kCensusK5];
    if (pattern == MGPM_P7)
      *count = 2 * (acc[kCensusEdgeTriples] - tek / 2 + 3 * kk - 10 * k5c);
    else
      *count = 3 * (acc[kCensusTriDegK] / 3 - tek + 6 * kk - 20 * k5c);
  }
  return MGPM_OK;
}

}  // namespace mgpm

// Degree-oriented DAG and small/large division built on the device.
//
// Reference counterpart: Graph::orientation_with_division (graph.cc:196-321)
// with parallel_prefix_sum (scan.h:4-35), which runs on the host.  Here
// every pass is flat over the arcs, so hub rows cost no more than any other
// arcs:
//   1. deg[v] (int32) and src[e] (the row of arc e, written row by row);
//   2. one value per arc: its class, as a count to be prefix-summed;
//   3. a device-wide exclusive scan of those values (hipCUB);
//   4. row pointers gathered from the scan at the row starts, and every
//      kept arc scattered to its scanned position (rows stay sorted).
// The division pass packs "small" into the low and "large" into the high
// 32 bits of one int64 per arc, so one scan gives both prefix counts.
//
// Classes, as the host builder (host_graph.cc build_dag / build_division):
//   DAG:       keep s->t iff d(t) > d(s) or (d(t) == d(s) and t > s);
//   division:  for nd = DAG out-degree, an arc {s,t} with min(nd) >= 1 is
//              "small" when max(nd) <= deg_thd and the DAG keeps s->t,
//              "large" when max(nd) > deg_thd and nd(s) > nd(t) or
//              (equal and s > t); otherwise neither.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "dag_device.hpp"
#include "wave.hpp"

namespace mgpm {

namespace {

constexpr int kDagBlock = 256;
constexpr int64_t kLarge = int64_t(1) << 32;  // unit of the "large" half

__global__ void vertex_degrees(const int64_t *__restrict__ rowptr, int32_t n,
                               int32_t *__restrict__ deg) {
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
       v += int64_t(gridDim.x) * blockDim.x)
    deg[v] = static_cast<int32_t>(rowptr[v + 1] - rowptr[v]);
}

// src[e] = v for every arc e of row v: one wave per row, coalesced stores.
__global__ __launch_bounds__(kDagBlock) void arc_sources(
    const int64_t *__restrict__ rowptr, int32_t n, int32_t *__restrict__ src) {
  const int lane = lane64();
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t v = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
       v < n; v += nwaves) {
    const int64_t e = rowptr[v + 1];
    for (int64_t i = rowptr[v] + lane; i < e; i += 64)
      src[i] = static_cast<int32_t>(v);
  }
}

__device__ __forceinline__ bool dag_keeps(const int32_t *__restrict__ deg,
                                          int32_t s, int32_t t) {
  const int32_t ds = deg[s], dt = deg[t];
  return dt > ds || (dt == ds && t > s);
}

// DIVISION: 1 (small) or kLarge (large) or 0; else 1 (kept) or 0.
template <bool DIVISION>
__global__ void arc_values(const int32_t *__restrict__ src,
                           const int32_t *__restrict__ col, int64_t nnz,
                           const int32_t *__restrict__ deg,
                           const int32_t *__restrict__ nd, int32_t thd,
                           int64_t *__restrict__ val) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int32_t s = src[e], t = col[e];
    int64_t x = 0;
    if (!DIVISION) {
      x = dag_keeps(deg, s, t) ? 1 : 0;
    } else {
      const int32_t a = nd[s], b = nd[t];
      if ((a < b ? a : b) >= 1) {
        if ((a > b ? a : b) <= thd) x = dag_keeps(deg, s, t) ? 1 : 0;
        else x = (a > b || (a == b && s > t)) ? kLarge : 0;
      }
    }
    val[e] = x;
  }
}

// Row pointers from the scanned values at the row starts; with DIVISION the
// two halves go to rowptr1 (small) and rowptr2 (large).  For the DAG pass
// also the out-degrees nd (int32).
template <bool DIVISION>
__global__ void row_offsets(const int64_t *__restrict__ rowptr, int32_t n,
                            const int64_t *__restrict__ pos,
                            int64_t *__restrict__ rowptr1,
                            int64_t *__restrict__ rowptr2,
                            int32_t *__restrict__ nd) {
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v <= n;
       v += int64_t(gridDim.x) * blockDim.x) {
    const int64_t p = pos[rowptr[v]];
    if (!DIVISION) {
      rowptr1[v] = p;
      if (v < n) nd[v] = static_cast<int32_t>(pos[rowptr[v + 1]] - p);
    } else {
      rowptr1[v] = p & (kLarge - 1);
      rowptr2[v] = p >> 32;
    }
  }
}

template <bool DIVISION>
__global__ void arc_scatter(const int32_t *__restrict__ col, int64_t nnz,
                            const int64_t *__restrict__ val,
                            const int64_t *__restrict__ pos,
                            int32_t *__restrict__ out1,
                            int32_t *__restrict__ out2) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = val[e];
    if (x == 0) continue;
    const int64_t p = pos[e];
    if (!DIVISION) out1[p] = col[e];
    else if (x == 1) out1[p & (kLarge - 1)] = col[e];
    else out2[p >> 32] = col[e];
  }
}

int flat_grid(int64_t items, int cus) {
  const int64_t want = (items + kDagBlock - 1) / kDagBlock;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, int64_t(cus) * 64)));
}

}  // namespace

int dag_scan_bytes(int64_t nnz, size_t *bytes) {
  size_t b = 0;
  const hipError_t e = hipcub::DeviceScan::ExclusiveSum(
      nullptr, b, static_cast<const int64_t *>(nullptr),
      static_cast<int64_t *>(nullptr), nnz + 1);
  *bytes = b;
  return e == hipSuccess ? 0 : -1;
}

int dag_device_build(const DagDevice &d, hipStream_t s) {
  const int gv = flat_grid(int64_t(d.n) + 1, d.cus);
  const int ga = flat_grid(d.nnz, d.cus);
  const int gw = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>((int64_t(d.n) + 3) / 4, int64_t(d.cus) * 32)));
  const size_t m1 = static_cast<size_t>(d.nnz) + 1;
  hipLaunchKernelGGL(vertex_degrees, dim3(gv), dim3(kDagBlock), 0, s,
                     d.rowptr, d.n, d.deg);
  hipLaunchKernelGGL(arc_sources, dim3(gw), dim3(kDagBlock), 0, s, d.rowptr,
                     d.n, d.src);
  // DAG.
  hipLaunchKernelGGL(arc_values<false>, dim3(ga), dim3(kDagBlock), 0, s,
                     d.src, d.col, d.nnz, d.deg, nullptr, 0, d.val);
  if (hipMemsetAsync(d.val + d.nnz, 0, sizeof(int64_t), s) != hipSuccess)
    return -1;
  size_t tb = d.scan_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(d.scan_tmp, tb, d.val, d.pos, m1, s) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(row_offsets<false>, dim3(gv), dim3(kDagBlock), 0, s,
                     d.rowptr, d.n, d.pos, d.out_rowptr[0], nullptr, d.nd);
  hipLaunchKernelGGL(arc_scatter<false>, dim3(ga), dim3(kDagBlock), 0, s,
                     d.col, d.nnz, d.val, d.pos, d.out_col[0], nullptr);
  if (!d.division) return hipGetLastError() == hipSuccess ? 0 : -1;
  // Division over the undirected adjacency, with the DAG out-degrees.
  hipLaunchKernelGGL(arc_values<true>, dim3(ga), dim3(kDagBlock), 0, s,
                     d.src, d.col, d.nnz, d.deg, d.nd, d.deg_thd, d.val);
  tb = d.scan_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(d.scan_tmp, tb, d.val, d.pos, m1, s) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(row_offsets<true>, dim3(gv), dim3(kDagBlock), 0, s,
                     d.rowptr, d.n, d.pos, d.out_rowptr[1], d.out_rowptr[2],
                     nullptr);
  hipLaunchKernelGGL(arc_scatter<true>, dim3(ga), dim3(kDagBlock), 0, s,
                     d.col, d.nnz, d.val, d.pos, d.out_col[1], d.out_col[2]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

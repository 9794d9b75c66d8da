// Batched multi-graph kernel: one persistent launch serves many (graph,
// pattern) queries.  Its queue holds work items from every query -- TC
// task groups, pattern task-edge groups, P1 vertex groups -- ordered by
// estimated cost across the whole batch, so large and small graphs share
// the device instead of running one launch (and one drain tail) each.
// The per-item work is the same device code as the per-graph kernels.
#include <hip/hip_runtime.h>

#include "batch_device.hpp"
#include "pattern_device_fns.hpp"
#include "tc_device_fns.hpp"

namespace mgpm {

namespace {

using namespace tcdev;
using namespace pdev;

constexpr int kBWaves = 4;
constexpr int kBLogp = 11;  // 2048-slot wave tables; larger ones binary-search
constexpr int kBItems = 8;

__device__ __forceinline__ void pattern_edges_item(const BatchQuery &q,
                                                   const PGraph &g, int t0,
                                                   int t1, Slab &s,
                                                   uint64_t &c1,
                                                   uint64_t &c2) {
  for (int t = t0; t < t1; ++t) {
    const int2 e = q.edges[t];
    switch (q.pattern) {
      case 1: edge_task<1>(g, e.x, e.y, s, c1, c2); break;
      case 2: edge_task<2>(g, e.x, e.y, s, c1, c2); break;
      case 3: edge_task<3>(g, e.x, e.y, s, c1, c2); break;
      case 4: edge_task<4>(g, e.x, e.y, s, c1, c2); break;
      case 5: edge_task<5>(g, e.x, e.y, s, c1, c2); break;
      case 6: edge_task<6>(g, e.x, e.y, s, c1, c2); break;
      case 7: edge_task<7>(g, e.x, e.y, s, c1, c2); break;
      case 8: edge_task<8>(g, e.x, e.y, s, c1, c2); break;
      default: break;
    }
  }
}

__global__ __launch_bounds__(kBWaves * 64) void batch_kernel(
    const BatchQuery *__restrict__ queries,
    const BatchItem *__restrict__ items, int nitems, unsigned *queue,
    unsigned long long *counters, int32_t *slab, int64_t cap) {
  __shared__ int32_t s_table[kBWaves][1 << kBLogp];
  __shared__ int32_t s_mark[kBWaves][64 * kBItems];
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  const int64_t wave = int64_t(blockIdx.x) * kBWaves + wv;
  Slab s(slab + wave * 8 * cap, cap);
  for (;;) {
    uint32_t w = 0;
    if (lane == 0) w = atomicAdd(queue, 1u);
    w = wave_first(w);
    if (w >= static_cast<uint32_t>(nitems)) break;
    const BatchItem it = items[w];
    const BatchQuery &q = queries[it.query];
    uint64_t c1 = 0, c2 = 0;
    if (it.kind == kItemTc) {
      const int32_t *goff = q.tc_groups[it.cls];
      const TcTask *tasks = q.tc_tasks[it.cls];
      const int t1 = goff[it.group + 1];
      for (int t = goff[it.group]; t < t1; ++t)
        c1 += wave_task<kBItems>(q.tc_col, q.tc_ncol, q.tc_partners, tasks[t],
                                 s_table[wv], kBLogp, s_mark[wv]);
    } else {
      const PGraph g{q.rowptr, q.col};
      if (it.kind == kItemEdges) {
        pattern_edges_item(q, g, q.edge_groups[it.group],
                           q.edge_groups[it.group + 1], s, c1, c2);
      } else {
        const int t1 = q.vert_groups[it.group + 1];
        for (int t = q.vert_groups[it.group]; t < t1; ++t)
          p1_vertex(g, q.verts[t], s, c1);
      }
    }
    c1 = wave_sum_u64(c1);
    c2 = wave_sum_u64(c2);
    if (lane == 0) {
      if (c1) atomicAdd(&counters[2 * it.query], (unsigned long long)c1);
      if (c2) atomicAdd(&counters[2 * it.query + 1], (unsigned long long)c2);
    }
  }
}

__global__ void batch_reset(unsigned *queue, unsigned long long *counters,
                            int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *queue = 0;
  if (i < n) counters[i] = 0;
}

}  // namespace

int batch_waves_per_block() { return kBWaves; }

int batch_device_launch(const BatchDevice &b, int grid, hipStream_t s) {
  const int n = 2 * b.nqueries;
  hipLaunchKernelGGL(batch_reset, dim3((n + 255) / 256 + 1), dim3(256), 0, s,
                     b.queue, b.counters, n);
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBWaves * 64), 0, s,
                     b.queries, b.items, b.nitems, b.queue, b.counters,
                     b.slab, b.cap);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

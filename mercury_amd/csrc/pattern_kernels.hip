// P1..P8 pattern-count kernels for gfx950.
//
// Reference counterpart: subgraph_matching/pattern_kernels.cuh:6-963 (one
// 32-lane warp per directed edge, frontier lists in a per-warp global slab,
// frequency arrays updated with atomics, 2-phase binary search).  The set
// algebra -- and therefore every count -- is the same; the machinery is
// wave64 throughout:
//   * one wave per task edge (P1 also per vertex), tasks taken from a
//     host-built cost-ordered queue of task groups instead of edge-id
//     striding, so hub edges start first and light edges are batched;
//   * heavy edges and vertices are split into shares, each its own queue
//     entry: a share runs the task's outer loops over every nsub-th index,
//     and the frequency factors the reference accumulates across the outer
//     loop (f[] of P2/P3/P5/P7/P8) are recounted per index instead, so one
//     hub edge is spread over many waves with no shared state;
//   * frontier lists are built with 64-lane ballot + mbcnt compaction;
//   * every "x in N(y)" test is an exact binary search, done by 64 lanes at
//     once over one uniform list N(y) (good locality), with the outer loop
//     over the shorter of two frontier lists;
//   * P1's unsplit vertex pass keeps its frequency array lane-owned
//     (index i always handled by lane i mod 64), so it needs no atomics;
//   * counts are exact uint64 everywhere (the reference narrows
//     warp_reduce results and their products to int32).
// Counters: counters[0] = the reference's counters[1] (c1), counters[1] =
// counters[2] (c2); sm_multigpu.cu:209-232 combines them on the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pattern_device_fns.hpp"

namespace mgpm {

namespace {

using namespace pdev;

constexpr int kPWaves = 4;

// One wave per task edge; queue entries are groups of task edges.
template <int P>
__global__ __launch_bounds__(kPWaves * 64) void pattern_edges(
    PGraph g, const int4 *__restrict__ edges,
    const int32_t *__restrict__ goff, int ngroups, int32_t *slab,
    int64_t cap, unsigned *queue, unsigned long long *counters) {
  const int lane = lane64();
  const int64_t wave = int64_t(blockIdx.x) * kPWaves + (threadIdx.x >> 6);
  Slab s(slab + wave * 8 * cap, cap);
  uint64_t c1 = 0, c2 = 0;
  for (;;) {
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(queue, 1u);
    q = wave_first(q);
    if (q >= static_cast<uint32_t>(ngroups)) break;
    const int t_end = goff[q + 1];
    for (int t = goff[q]; t < t_end; ++t) {
      const int4 e = edges[t];
      edge_task<P>(g, e.x, e.y, e.z, e.w, s, c1, c2);
    }
  }
  flush2(c1, c2, counters);
}

// P1's vertex pass (frequency count over every vertex).
__global__ __launch_bounds__(kPWaves * 64) void pattern_p1_vertices(
    PGraph g, const int4 *__restrict__ verts,
    const int32_t *__restrict__ goff, int ngroups, int32_t *slab,
    int64_t cap, unsigned *queue, unsigned long long *counters) {
  const int lane = lane64();
  const int64_t wave = int64_t(blockIdx.x) * kPWaves + (threadIdx.x >> 6);
  Slab s(slab + wave * 8 * cap, cap);
  uint64_t c1 = 0, c2 = 0;
  for (;;) {
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(queue, 1u);
    q = wave_first(q);
    if (q >= static_cast<uint32_t>(ngroups)) break;
    const int t_end = goff[q + 1];
    for (int t = goff[q]; t < t_end; ++t) {
      const int4 v = verts[t];
      p1_vertex(g, v.x, v.z, v.w, s, c1);
    }
  }
  flush2(c1, c2, counters);
}

__global__ void pattern_reset(unsigned *queues, unsigned long long *counters) {
  const int i = threadIdx.x;
  if (i < 2) queues[i] = 0;
  if (i < 2) counters[i] = 0;
}

template <int P>
void launch_edges(const PatternDevicePlan &p, int grid, hipStream_t s) {
  PGraph g{p.rowptr, p.col};
  hipLaunchKernelGGL(pattern_edges<P>, dim3(grid), dim3(kPWaves * 64), 0, s,
                     g, p.edges, p.edge_groups, p.n_edge_groups, p.slab,
                     p.cap, p.queues, p.counters);
}

// Workgroups that can take a queue entry: one entry per wave.
int blocks_for(int groups) { return (groups + kPWaves - 1) / kPWaves; }

}  // namespace

int pattern_waves_per_block() { return kPWaves; }

int pattern_device_launch(const PatternDevicePlan &p, int pattern, int grid,
                          hipStream_t s, bool reset) {
  if (pattern < 1 || pattern > 8 || grid < 1) return -1;
  if (reset)
    hipLaunchKernelGGL(pattern_reset, dim3(1), dim3(64), 0, s, p.queues,
                       p.counters);
  if (pattern == 1 && p.n_vert_groups > 0) {
    PGraph g{p.rowptr, p.col};
    const int vgrid = std::min(grid, blocks_for(p.n_vert_groups));
    hipLaunchKernelGGL(pattern_p1_vertices, dim3(vgrid), dim3(kPWaves * 64), 0,
                       s, g, p.verts, p.vert_groups, p.n_vert_groups, p.slab,
                       p.cap, p.queues + 1, p.counters);
  }
  if (p.n_edge_groups <= 0) return hipGetLastError() == hipSuccess ? 0 : -1;
  grid = std::min(grid, blocks_for(p.n_edge_groups));
  switch (pattern) {
    case 1: launch_edges<1>(p, grid, s); break;
    case 2: launch_edges<2>(p, grid, s); break;
    case 3: launch_edges<3>(p, grid, s); break;
    case 4: launch_edges<4>(p, grid, s); break;
    case 5: launch_edges<5>(p, grid, s); break;
    case 6: launch_edges<6>(p, grid, s); break;
    case 7: launch_edges<7>(p, grid, s); break;
    case 8: launch_edges<8>(p, grid, s); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

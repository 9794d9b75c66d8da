// Triangle-count plan built on the device (the host builder is
// tc_plan.cc; same owner rule, same task classes, same cost model).
//
// Every step is a flat pass over arcs, partners or tasks, plus hipCUB
// scans and one radix sort:
//   1. owned(e) for every undirected arc e = (t, x): the endpoint with the
//      larger DAG out-degree owns the edge (ties: larger id), edges with an
//      empty side are dropped -- scan -> partner index;
//   2. partners scattered in CSR order: {beg = DAG row of x, len = d+(x)},
//      with their owner and cost (len + kTcPartnerCost);
//   3. cost prefix (scan); a partner starts a task where its owner's list
//      starts or where the owner-relative cost prefix crosses a multiple of
//      chunk_work (tasks stay below chunk_work plus one partner, like the
//      host's greedy cut) -- scan -> task index;
//   4. task records and their cost (table + list + partners);
//   5. sort by cost, descending and stable (the host's LPT order, ties by
//      task index);
//   6. stable partition into the four table classes (two packed scans);
//   7. queue groups per class: a group ends where the class-relative cost
//      prefix crosses a multiple of the class limit (as the host:
//      group_work, lowered to class cost / kTcTargetGroups, at least
//      kTcMinGroupWork) and at most every 64 tasks.
// Scalar results (counts, class sizes and limits, S+, max d+) collect in a
// small device summary that the host reads once at the end.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "tc_dplan.hpp"
#include "wave.hpp"

namespace mgpm {

namespace {

constexpr int kBlock = 256;

int flat_grid(int64_t items, int cus) {
  const int64_t want = (items + kBlock - 1) / kBlock;
  return static_cast<int>(
      std::max<int64_t>(1, std::min<int64_t>(want, int64_t(cus) * 64)));
}

#define MGPM_GRID_LOOP(i, count)                                       \
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;     \
       i < (count); i += int64_t(gridDim.x) * blockDim.x)

__device__ __forceinline__ bool owns(const int32_t *__restrict__ nd, int32_t t,
                                     int32_t x) {
  const int32_t a = nd[t], b = nd[x];
  return a >= 1 && b >= 1 && (a > b || (a == b && t > x));
}

__global__ void owned_values(const int32_t *__restrict__ src,
                             const int32_t *__restrict__ col, int64_t nnz,
                             const int32_t *__restrict__ nd,
                             int64_t *__restrict__ val) {
  MGPM_GRID_LOOP(e, nnz) val[e] = owns(nd, src[e], col[e]) ? 1 : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) val[nnz] = 0;
}

__global__ void partner_scatter(const int32_t *__restrict__ src,
                                const int32_t *__restrict__ col, int64_t nnz,
                                const int64_t *__restrict__ val,
                                const int64_t *__restrict__ pos,
                                const int64_t *__restrict__ dag_rowptr,
                                const int32_t *__restrict__ nd,
                                TcPartner *__restrict__ partners,
                                int32_t *__restrict__ owner,
                                int64_t *__restrict__ cost) {
  MGPM_GRID_LOOP(e, nnz) {
    if (!val[e]) continue;
    const int64_t p = pos[e];
    const int32_t x = col[e];
    TcPartner pr;
    pr.beg = static_cast<uint32_t>(dag_rowptr[x]);
    pr.len = static_cast<uint32_t>(nd[x]);
    partners[p] = pr;
    owner[p] = src[e];
    cost[p] = int64_t(pr.len) + kTcPartnerCost;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cost[pos[nnz]] = 0;
}

// flag[i] = 1 where partner i starts a task.  `first` of owner t is
// pos[rowptr[t]], the index of its first partner.
__global__ void task_flags(const int32_t *__restrict__ owner,
                           const int64_t *__restrict__ cpre, int64_t np,
                           const int64_t *__restrict__ rowptr,
                           const int64_t *__restrict__ pos, int64_t chunk,
                           int64_t *__restrict__ flag) {
  MGPM_GRID_LOOP(i, np) {
    const int64_t first = pos[rowptr[owner[i]]];
    int64_t f = 1;
    if (i != first) {
      const int64_t base = cpre[first];
      f = (cpre[i] - base) / chunk != (cpre[i - 1] - base) / chunk ? 1 : 0;
    }
    flag[i] = f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[np] = 0;
}

__global__ void task_starts(const int64_t *__restrict__ flag,
                            const int64_t *__restrict__ tix, int64_t np,
                            int32_t *__restrict__ start) {
  MGPM_GRID_LOOP(i, np) if (flag[i]) start[tix[i]] = static_cast<int32_t>(i);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    start[tix[np]] = static_cast<int32_t>(np);
}

__global__ void task_records(const int32_t *__restrict__ start, int64_t nt,
                             const int32_t *__restrict__ owner,
                             const int64_t *__restrict__ cpre,
                             const int64_t *__restrict__ dag_rowptr,
                             const int32_t *__restrict__ nd,
                             TcTask *__restrict__ tasks,
                             uint64_t *__restrict__ keys,
                             int32_t *__restrict__ vals) {
  MGPM_GRID_LOOP(k, nt) {
    const int32_t s = start[k], e = start[k + 1];
    const int32_t t = owner[s];
    TcTask task;
    task.tbeg = static_cast<uint32_t>(dag_rowptr[t]);
    task.tlen = nd[t];
    task.pbeg = static_cast<uint32_t>(s);
    task.pcnt = e - s;
    tasks[k] = task;
    keys[k] = static_cast<uint64_t>(kTcTableCost + task.tlen + cpre[e] -
                                    cpre[s]);
    vals[k] = static_cast<int32_t>(k);
  }
}

__device__ __forceinline__ int task_class(int32_t tlen, const TcDplanClasses &c) {
  if (tlen <= c.small_max) return kTcSmall;
  if (tlen <= c.wave_max) return kTcMedium;
  if (tlen <= c.block_max) return kTcBlock;
  return kTcGlobal;
}

// Class membership as two packed counters per sorted task (classes 0/1 in
// the low/high half of a, 2/3 in b), class cost and largest table.
__global__ void class_values(const int32_t *__restrict__ order,
                             const uint64_t *__restrict__ keys, int64_t nt,
                             const TcTask *__restrict__ tasks,
                             TcDplanClasses cls, int64_t *__restrict__ a,
                             int64_t *__restrict__ b,
                             TcDplanSummary *__restrict__ sum) {
  uint64_t work[kTcNumClasses] = {0, 0, 0, 0};
  int32_t maxlen[kTcNumClasses] = {0, 0, 0, 0};
  MGPM_GRID_LOOP(j, nt) {
    const int32_t tlen = tasks[order[j]].tlen;
    const int c = task_class(tlen, cls);
    const int64_t one = int64_t(1) << (32 * (c & 1));
    a[j] = c < 2 ? one : 0;
    b[j] = c < 2 ? 0 : one;
#pragma unroll
    for (int cc = 0; cc < kTcNumClasses; ++cc) {  // (no dynamic indexing)
      if (cc != c) continue;
      work[cc] += keys[j];
      maxlen[cc] = maxlen[cc] > tlen ? maxlen[cc] : tlen;
    }
  }
  // One atomic per wave and class (the whole wave is converged here).
#pragma unroll
  for (int c = 0; c < kTcNumClasses; ++c) {
    const unsigned long long w = wave_sum_u64(work[c]);
    const int32_t mx = wave_max_i32(maxlen[c]);
    if (lane64() == 0) {
      if (w) atomicAdd(&sum->class_work[c], w);
      if (mx) atomicMax(&sum->class_max_tlen[c], mx);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a[nt] = 0;
    b[nt] = 0;
  }
}

// One thread: class sizes, offsets and group limits.
__global__ void class_setup(const int64_t *__restrict__ as,
                            const int64_t *__restrict__ bs, int64_t nt,
                            int64_t group_work,
                            TcDplanSummary *__restrict__ sum) {
  const int64_t mask = (int64_t(1) << 32) - 1;
  sum->class_cnt[0] = as[nt] & mask;
  sum->class_cnt[1] = as[nt] >> 32;
  sum->class_cnt[2] = bs[nt] & mask;
  sum->class_cnt[3] = bs[nt] >> 32;
  int64_t off = 0;
  for (int c = 0; c < kTcNumClasses; ++c) {
    sum->class_off[c] = off;
    off += sum->class_cnt[c];
    int64_t lim = static_cast<int64_t>(sum->class_work[c]) / kTcTargetGroups;
    lim = lim < group_work ? lim : group_work;
    sum->class_limit[c] = lim > kTcMinGroupWork ? lim : kTcMinGroupWork;
  }
}

__global__ void class_scatter(const int32_t *__restrict__ order,
                              const uint64_t *__restrict__ keys, int64_t nt,
                              const TcTask *__restrict__ tasks,
                              TcDplanClasses cls,
                              const int64_t *__restrict__ as,
                              const int64_t *__restrict__ bs,
                              const TcDplanSummary *__restrict__ sum,
                              TcTask *__restrict__ out,
                              int64_t *__restrict__ work) {
  const int64_t mask = (int64_t(1) << 32) - 1;
  MGPM_GRID_LOOP(j, nt) {
    const TcTask task = tasks[order[j]];
    const int c = task_class(task.tlen, cls);
    const int64_t packed = c < 2 ? as[j] : bs[j];
    const int64_t rank = (c & 1) ? (packed >> 32) : (packed & mask);
    const int64_t d = sum->class_off[c] + rank;
    out[d] = task;
    work[d] = static_cast<int64_t>(keys[j]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) work[nt] = 0;
}

__device__ __forceinline__ int class_of_slot(int64_t j,
                                             const TcDplanSummary *sum) {
  int c = 0;
  while (c + 1 < kTcNumClasses && j >= sum->class_off[c + 1]) ++c;
  return c;
}

// Cost boundaries of the queue groups: a class's first task, and every
// task at which the class's work prefix crosses a multiple of its limit.
__global__ void group_flags(const int64_t *__restrict__ wpre, int64_t nt,
                            const TcDplanSummary *__restrict__ sum,
                            int64_t *__restrict__ flag) {
  MGPM_GRID_LOOP(j, nt) {
    const int c = class_of_slot(j, sum);
    const int64_t off = sum->class_off[c];
    const int64_t r = j - off;
    int64_t f = 1;
    if (r != 0) {
      const int64_t base = wpre[off], lim = sum->class_limit[c];
      f = (wpre[j] - base) / lim != (wpre[j - 1] - base) / lim ? 1 : 0;
    }
    flag[j] = f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[nt] = 0;
}

// First task of every cost group (gix: exclusive scan of the cost flags).
__global__ void group_starts(const int64_t *__restrict__ flag,
                             const int64_t *__restrict__ gix, int64_t nt,
                             int64_t *__restrict__ start) {
  MGPM_GRID_LOOP(j, nt) {
    if (flag[j]) start[gix[j]] = j;
  }
}

// Final boundaries: the cost boundaries, and every 64th task counted from
// the start of its cost group, so a queue entry holds at most 64 tasks and
// light tasks fill it (the host builder's greedy cut closes a group at 64
// tasks the same way).
__global__ void group_cap_flags(const int64_t *__restrict__ flag,
                                const int64_t *__restrict__ gix,
                                const int64_t *__restrict__ start, int64_t nt,
                                int64_t *__restrict__ out) {
  MGPM_GRID_LOOP(j, nt) {
    const int64_t g = gix[j + 1] - 1;  // j's cost group
    out[j] = (flag[j] || (j - start[g]) % 64 == 0) ? 1 : 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[nt] = 0;
}

// Group offsets of class c live at groups + class_off[c] + c (its tasks'
// count plus one entries at most).
__global__ void group_offsets(const int64_t *__restrict__ flag,
                              const int64_t *__restrict__ gix, int64_t nt,
                              TcDplanSummary *__restrict__ sum,
                              int32_t *__restrict__ groups) {
  MGPM_GRID_LOOP(j, nt) {
    if (!flag[j]) continue;
    const int c = class_of_slot(j, sum);
    const int64_t off = sum->class_off[c];
    groups[off + c + (gix[j] - gix[off])] = static_cast<int32_t>(j - off);
  }
  if (blockIdx.x == 0 && threadIdx.x < kTcNumClasses) {
    const int c = threadIdx.x;
    const int64_t off = sum->class_off[c], cnt = sum->class_cnt[c];
    const int64_t ng = gix[off + cnt] - gix[off];
    sum->ngroups[c] = ng;
    groups[off + c + ng] = static_cast<int32_t>(cnt);
  }
}

// S+ = sum over DAG arcs u->v of d+(v); largest d+.
__global__ void dag_sums(const int32_t *__restrict__ dcol, int64_t mplus,
                         const int32_t *__restrict__ nd, int32_t n,
                         TcDplanSummary *__restrict__ sum) {
  unsigned long long s = 0;
  int32_t mx = 0;
  MGPM_GRID_LOOP(i, mplus) s += static_cast<unsigned long long>(nd[dcol[i]]);
  MGPM_GRID_LOOP(v, n) mx = nd[v] > mx ? nd[v] : mx;
  s = wave_sum_u64(s);
  mx = wave_max_i32(mx);
  if (lane64() == 0) {
    if (s) atomicAdd(&sum->s_plus, s);
    if (mx) atomicMax(&sum->max_out, mx);
  }
}

}  // namespace

int dplan_scan_bytes(int64_t count, size_t *bytes) {
  size_t b = 0;
  const hipError_t e = hipcub::DeviceScan::ExclusiveSum(
      nullptr, b, static_cast<const int64_t *>(nullptr),
      static_cast<int64_t *>(nullptr), count);
  *bytes = b;
  return e == hipSuccess ? 0 : -1;
}

int dplan_scan(void *tmp, size_t bytes, const int64_t *in, int64_t *out,
               int64_t count, hipStream_t s) {
  size_t b = bytes;
  return hipcub::DeviceScan::ExclusiveSum(tmp, b, in, out, count, s) ==
                 hipSuccess
             ? 0
             : -1;
}

int dplan_sort_bytes(int64_t count, size_t *bytes) {
  size_t b = 0;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(
      nullptr, b, static_cast<const uint64_t *>(nullptr),
      static_cast<uint64_t *>(nullptr), static_cast<const int32_t *>(nullptr),
      static_cast<int32_t *>(nullptr), static_cast<int>(count));
  *bytes = b;
  return e == hipSuccess ? 0 : -1;
}

int dplan_sort(void *tmp, size_t bytes, const uint64_t *keys_in,
               uint64_t *keys_out, const int32_t *vals_in, int32_t *vals_out,
               int64_t count, hipStream_t s) {
  size_t b = bytes;
  return hipcub::DeviceRadixSort::SortPairsDescending(
             tmp, b, keys_in, keys_out, vals_in, vals_out,
             static_cast<int>(count), 0, 64, s) == hipSuccess
             ? 0
             : -1;
}

int dplan_partners(const TcDplanArgs &a, hipStream_t s) {
  const int g = flat_grid(a.nnz, a.cus);
  hipLaunchKernelGGL(owned_values, dim3(g), dim3(kBlock), 0, s, a.src, a.col,
                     a.nnz, a.nd, a.val);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.val, a.pos, a.nnz + 1, s))
    return -1;
  hipLaunchKernelGGL(partner_scatter, dim3(g), dim3(kBlock), 0, s, a.src,
                     a.col, a.nnz, a.val, a.pos, a.dag_rowptr, a.nd,
                     a.partners, a.owner, a.cost);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dplan_task_flags(const TcDplanArgs &a, int64_t np, hipStream_t s) {
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.cost, a.cpre, np + 1, s))
    return -1;
  hipLaunchKernelGGL(task_flags, dim3(flat_grid(np, a.cus)), dim3(kBlock), 0,
                     s, a.owner, a.cpre, np, a.rowptr, a.pos, a.chunk_work,
                     a.val);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.val, a.tix, np + 1, s))
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dplan_tasks(const TcDplanArgs &a, int64_t np, int64_t nt, hipStream_t s) {
  hipLaunchKernelGGL(task_starts, dim3(flat_grid(np, a.cus)), dim3(kBlock), 0,
                     s, a.val, a.tix, np, a.start);
  const int g = flat_grid(nt, a.cus);
  hipLaunchKernelGGL(task_records, dim3(g), dim3(kBlock), 0, s, a.start, nt,
                     a.owner, a.cpre, a.dag_rowptr, a.nd, a.tasks_all, a.keys,
                     a.vals);
  if (dplan_sort(a.sort_tmp, a.sort_bytes, a.keys, a.keys_sorted, a.vals,
                 a.order, nt, s))
    return -1;
  hipLaunchKernelGGL(class_values, dim3(g), dim3(kBlock), 0, s, a.order,
                     a.keys_sorted, nt, a.tasks_all, a.cls, a.ca, a.cb,
                     a.summary);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.ca, a.cas, nt + 1, s)) return -1;
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.cb, a.cbs, nt + 1, s)) return -1;
  hipLaunchKernelGGL(class_setup, dim3(1), dim3(1), 0, s, a.cas, a.cbs, nt,
                     a.group_work, a.summary);
  hipLaunchKernelGGL(class_scatter, dim3(g), dim3(kBlock), 0, s, a.order,
                     a.keys_sorted, nt, a.tasks_all, a.cls, a.cas, a.cbs,
                     a.summary, a.tasks_out, a.work);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.work, a.wpre, nt + 1, s))
    return -1;
  hipLaunchKernelGGL(group_flags, dim3(g), dim3(kBlock), 0, s, a.wpre, nt,
                     a.summary, a.gflag);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.gflag, a.gix, nt + 1, s))
    return -1;
  // ca, cb, cas are dead after class_scatter: cost-group starts, final
  // flags and their scan.
  hipLaunchKernelGGL(group_starts, dim3(g), dim3(kBlock), 0, s, a.gflag,
                     a.gix, nt, a.ca);
  hipLaunchKernelGGL(group_cap_flags, dim3(g), dim3(kBlock), 0, s, a.gflag,
                     a.gix, a.ca, nt, a.cb);
  if (dplan_scan(a.scan_tmp, a.scan_bytes, a.cb, a.cas, nt + 1, s))
    return -1;
  hipLaunchKernelGGL(group_offsets, dim3(g), dim3(kBlock), 0, s, a.cb,
                     a.cas, nt, a.summary, a.groups_out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dplan_dag_sums(const int32_t *dcol, int64_t mplus, const int32_t *nd,
                   int32_t n, int cus, TcDplanSummary *sum, hipStream_t s) {
  hipLaunchKernelGGL(dag_sums, dim3(flat_grid(std::max<int64_t>(mplus, n), cus)),
                     dim3(kBlock), 0, s, dcol, mplus, nd, n, sum);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

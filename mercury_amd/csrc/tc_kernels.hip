// Triangle-count kernels for gfx950.
//
// Reference counterpart: gpu_kernels/tc_challenge.cuh:4-303
// (hashIndex_for_large_degree, binarySearch_for_small_degree).  Same sum
// (|N+(a) ∩ N+(b)| over DAG edges), different machinery:
//   * every edge hashes the larger DAG out-list and streams the smaller one
//     (the reference does this only for its "large" edges);
//   * tables live in LDS as open-addressing hash sets sized to the task
//     (load factor <= 1/2), wave-private up to 2048 slots, workgroup-shared
//     up to 32768 slots, with a global binary-search path beyond that (the
//     reference's 4096-entry LDS table has no such path);
//   * partner lists are flattened across all 64 lanes, 4 items per lane per
//     step (segment owners found with ballot + mbcnt over LDS start marks),
//     so short lists do not idle lanes and each wave keeps 4 independent
//     gathers in flight;
//   * one persistent dynamic queue per task class replaces the
//     BLOCK/WARP_DYNAMIC_CATCH counters (tuning_schedules.cuh:3-13); its
//     entries are host-built groups of tasks of bounded cost (heavy tasks
//     alone, light ones packed), which keeps the queue atomic off the
//     critical path without putting several heavy tasks on one wave.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "tc_device_fns.hpp"

namespace mgpm {

namespace {

using namespace tcdev;

constexpr int kSmallLogpHalf = 8;  // half-wave tables: 256 slots

// Wave-private tables of up to 2^LOGP_MAX slots (tasks with 2*d+ <= that).
// STREAM: every task streams its laid-out items (items[item_off[pbeg] ..
// item_off[pbeg + pcnt])) instead of gathering its partner lists.
template <int LOGP_MAX, int WAVES, int ITEMS, bool STREAM>
__global__ __launch_bounds__(WAVES * 64) void tc_wave_tables(
    const int32_t *__restrict__ col, uint32_t ncol,
    const TcTask *__restrict__ tasks, const int32_t *__restrict__ goff,
    int ngroups, const TcPartner *__restrict__ partners,
    const int32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
    unsigned *__restrict__ queue, unsigned long long *__restrict__ result,
    int cursor_min, int merge_max) {
  __shared__ int32_t s_table[WAVES][1 << LOGP_MAX];
  __shared__ int32_t s_mark[WAVES][STREAM ? 1 : 64 * ITEMS];  // start marks
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  int32_t *table = s_table[wv];
  unsigned long long acc = 0;
  for (;;) {
    uint32_t g = 0;
    if (lane == 0) g = atomicAdd(queue, 1u);
    g = wave_first(g);
    if (g >= static_cast<uint32_t>(ngroups)) break;
    const int t_end = goff[g + 1];
    int t = goff[g];
    if constexpr (STREAM) {
      // A three-stage pipeline over the group's tasks, so that a task's
      // loads are all issued in earlier iterations: iteration t issues the
      // record of task t + 3, the item offsets and owner items of t + 2
      // (from its record, loaded one iteration earlier) and the first step
      // of items of t + 1 (from its offsets, loaded one iteration earlier).
      // Indices past the group are clamped to its last task.
      const int last = t_end - 1;
      TcTask r0 = tasks[t];
      TcTask r1 = tasks[t + 1 < last ? t + 1 : last];
      TcTask r2 = tasks[t + 2 < last ? t + 2 : last];
      TaskPre p0 = load_task_pre<true>(col, partners, r0, lane, item_off);
      TaskPre p1 = load_task_pre<true>(col, partners, r1, lane, item_off);
      FirstStep<ITEMS> f0 = load_first_step<ITEMS>(items, p0.ib, p0.ie, lane);
      for (; t < t_end; ++t) {
        const TcTask r3 = tasks[t + 3 < last ? t + 3 : last];
        const TaskPre p2 =
            load_task_pre<true>(col, partners, r2, lane, item_off);
        const FirstStep<ITEMS> f1 =
            load_first_step<ITEMS>(items, p1.ib, p1.ie, lane);
        if (r0.tlen <= merge_max)
          acc += wave_task_merge(col, partners, items, item_off, r0, table,
                                 p0);
        else
          acc += wave_task_stream<ITEMS>(col, items, r0, table, LOGP_MAX, p0,
                                         &f0);
        r0 = r1;
        r1 = r2;
        r2 = r3;
        p0 = p1;
        p1 = p2;
        f0 = f1;
      }
      continue;
    }
    TcTask task = tasks[t];
    TaskPre pre = load_task_pre<STREAM>(col, partners, task, lane, item_off);
    for (; t < t_end; ++t) {
      // The next task's loads go out before this task's gathers (the last
      // task of the group reloads itself: same load count on every path).
      const TcTask next = tasks[t + 1 < t_end ? t + 1 : t];
      const TaskPre next_pre =
          load_task_pre<STREAM>(col, partners, next, lane, item_off);
      if (task.tlen <= merge_max)
        acc += wave_task_merge(col, partners, STREAM ? items : nullptr,
                               item_off, task, table, pre);
      else if constexpr (STREAM)
        acc += wave_task_stream<ITEMS>(col, items, task, table, LOGP_MAX,
                                       pre);
      else
        acc += wave_task<ITEMS>(col, ncol, partners, task, table, LOGP_MAX,
                                s_mark[wv], cursor_min, pre);
      task = next;
      pre = next_pre;
    }
  }
  flush_hits(acc, result);
}

// One table per workgroup (dynamic LDS of 2^logp_max slots).
template <int WAVES, int ITEMS, bool STREAM>
__global__ __launch_bounds__(WAVES * 64) void tc_block_tables(
    const int32_t *__restrict__ col, uint32_t ncol,
    const TcTask *__restrict__ tasks, const int32_t *__restrict__ goff,
    int ngroups, const TcPartner *__restrict__ partners,
    const int32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
    unsigned *__restrict__ queue, unsigned long long *__restrict__ result,
    int logp_max, int cursor_min) {
  extern __shared__ int32_t s_table[];
  __shared__ int32_t s_mark[WAVES][STREAM ? 1 : 64 * ITEMS];  // start marks
  __shared__ uint32_t s_task;
  const int wv = threadIdx.x >> 6;
  unsigned long long acc = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t g = s_task;
    if (g >= static_cast<uint32_t>(ngroups)) break;
    const int t_end = goff[g + 1];
    for (int t = goff[g]; t < t_end; ++t) {
    const TcTask task = tasks[t];
    uint64_t ib = 0, ie = 0;
    if constexpr (STREAM) {
      ib = item_off[task.pbeg];
      ie = item_off[task.pbeg + static_cast<uint32_t>(task.pcnt)];
    }
    if (2 * task.tlen > (1 << logp_max)) {  // never planned; kept exact
      SortedProbe probe{col + task.tbeg, task.tlen};
      if constexpr (STREAM)
        acc += stream_items<ITEMS>(items, ib, ie, probe, wv, WAVES);
      else
        acc += stream_partners<ITEMS>(col, ncol, partners, task.pbeg,
                                      task.pcnt, wv, WAVES, s_mark[wv], probe,
                                      cursor_min);
      __syncthreads();
      continue;
    }
    int logp = ceil_log2(2 * task.tlen);
    logp = logp < 6 ? 6 : (logp > logp_max ? logp_max : logp);
    const int P = 1 << logp;
    for (int i = threadIdx.x; i < P; i += WAVES * 64) s_table[i] = -1;
    __syncthreads();
    table_insert(s_table, logp, col, task.tbeg, task.tlen, threadIdx.x,
                 WAVES * 64, true);
    __syncthreads();
    HashProbe<true> probe{s_table, logp, static_cast<uint32_t>(P - 1)};
    if constexpr (STREAM)
      acc += stream_items<ITEMS>(items, ib, ie, probe, wv, WAVES);
    else
      acc += stream_partners<ITEMS>(col, ncol, partners, task.pbeg, task.pcnt,
                                    wv, WAVES, s_mark[wv], probe, cursor_min);
    __syncthreads();
    }
  }
  flush_hits(acc, result);
}

// Tables too large for LDS: binary search in the global sorted out-list.
template <int WAVES, int ITEMS, bool STREAM>
__global__ __launch_bounds__(WAVES * 64) void tc_global_search(
    const int32_t *__restrict__ col, uint32_t ncol,
    const TcTask *__restrict__ tasks, const int32_t *__restrict__ goff,
    int ngroups, const TcPartner *__restrict__ partners,
    const int32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
    unsigned *__restrict__ queue, unsigned long long *__restrict__ result,
    int cursor_min) {
  __shared__ int32_t s_mark[WAVES][STREAM ? 1 : 64 * ITEMS];  // start marks
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  unsigned long long acc = 0;
  for (;;) {
    uint32_t g = 0;
    if (lane == 0) g = atomicAdd(queue, 1u);
    g = wave_first(g);
    if (g >= static_cast<uint32_t>(ngroups)) break;
    const int t_end = goff[g + 1];
    for (int t = goff[g]; t < t_end; ++t) {
      const TcTask task = tasks[t];
      SortedProbe probe{col + task.tbeg, task.tlen};
      if constexpr (STREAM)
        acc += stream_items<ITEMS>(
            items, item_off[task.pbeg],
            item_off[task.pbeg + static_cast<uint32_t>(task.pcnt)], probe, 0,
            1);
      else
        acc += stream_partners<ITEMS>(col, ncol, partners, task.pbeg,
                                      task.pcnt, 0, 1, s_mark[wv], probe,
                                      cursor_min);
    }
  }
  flush_hits(acc, result);
}

constexpr int kOwnPerLane = kTcSmallMaxDeg / 32;  // owner items per lane

// Small-class tasks on laid-out lists, two per wave: lanes 0-31 take the
// group's tasks t0, t0 + 2, ... and lanes 32-63 t0 + 1, t0 + 3, ..., each
// half with its own table (up to 256 slots) and its own items, 32 x ITEMS
// per step; the halves step together (a half without a task, or with
// fewer steps, runs idle items).  Tiny tasks (ER-like graphs: about 16
// owner items and 170 streamed items) then share each wave's per-task
// costs -- table clear and fill, synchronisation, the load chain.
template <int WAVES, int ITEMS>
__global__ __launch_bounds__(WAVES * 64) void tc_half_tables(
    const int32_t *__restrict__ col, const TcTask *__restrict__ tasks,
    const int32_t *__restrict__ goff, int ngroups,
    const int32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
    unsigned *__restrict__ queue, unsigned long long *__restrict__ result) {
  __shared__ int32_t s_table[WAVES][2][1 << kSmallLogpHalf];
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  const int sub = lane >> 5, gl = lane & 31;
  int32_t *table = s_table[wv][sub];
  unsigned long long acc = 0;
  for (;;) {
    uint32_t g = 0;
    if (lane == 0) g = atomicAdd(queue, 1u);
    g = wave_first(g);
    if (g >= static_cast<uint32_t>(ngroups)) break;
    const int t0 = goff[g], t_end = goff[g + 1];
    // Loads run ahead so that a pair waits about one memory round trip:
    // each pair issues the record two pairs ahead, the next pair's item
    // offsets (its record came in during the previous pair), its own owner
    // items and its first two steps of streamed items together, at
    // clamped (valid) addresses, before it clears and fills its table.  An
    // idle half reads the other half's task.
    auto rec_of = [&](int t) {
      const int tt = t + sub;
      return tasks[tt < t_end ? tt : (t < t_end ? t : t_end - 1)];
    };
    TcTask task = rec_of(t0);
    TcTask next = rec_of(t0 + 2);
    uint64_t ib = item_off[task.pbeg];
    uint64_t ie = item_off[task.pbeg + static_cast<uint32_t>(task.pcnt)];
    for (int t = t0; t < t_end; t += 2) {
      const bool have = t + sub < t_end;
      const TcTask next2 = rec_of(t + 4);
      const uint64_t nib = item_off[next.pbeg];
      const uint64_t nie =
          item_off[next.pbeg + static_cast<uint32_t>(next.pcnt)];
      const int tl = task.tlen;  // 1 .. kTcSmallMaxDeg
      int32_t own[kOwnPerLane];
#pragma unroll
      for (int j = 0; j < kOwnPerLane; ++j) {
        const int k = gl + 32 * j;
        own[j] = col[task.tbeg + static_cast<uint32_t>(k < tl ? k : 0)];
      }
      const uint64_t n = have ? ie - ib : 0;
      int32_t w[ITEMS], w1[ITEMS];
      bool ok[ITEMS], ok1[ITEMS];
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {  // the first two steps
        const uint64_t k = uint64_t(32 * i + gl);
        ok[i] = k < n;
        ok1[i] = k + 32 * ITEMS < n;
        const int32_t v = items[ib + (ok[i] ? k : 0)];  // clamped: ib < ie
        const int32_t v1 = items[ib + (ok1[i] ? k + 32 * ITEMS : 0)];
        w[i] = ok[i] ? v : -2;
        w1[i] = ok1[i] ? v1 : -2;
      }
      int logp = ceil_log2(8 * tl);
      logp = logp < 5 ? 5 : (logp > kSmallLogpHalf ? kSmallLogpHalf : logp);
      const int P = 1 << logp;
      for (int i = gl; i < P; i += 32) table[i] = -1;
      wave_sync();
      if (have) {
        const uint32_t mask = static_cast<uint32_t>(P - 1);
#pragma unroll
        for (int j = 0; j < kOwnPerLane; ++j) {
          if (gl + 32 * j >= tl) continue;
          const int32_t key = own[j];
          uint32_t h = slot_of(key, logp);
          while (atomicCAS(&table[h], -1, key) != -1) h = (h + 1) & mask;
        }
      }
      wave_sync();
      HashProbe<false> probe{table, logp, static_cast<uint32_t>(P - 1)};
      acc += probe.template lookup<ITEMS>(w, ok);  // the first two steps
      acc += probe.template lookup<ITEMS>(w1, ok1);
      const uint64_t steps = (n + 32 * ITEMS - 1) / (32 * ITEMS);
      const uint64_t other = __shfl_xor(steps, 32, 64);
      const uint64_t wsteps = steps > other ? steps : other;
      for (uint64_t st = 2; st < wsteps; ++st) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
          const uint64_t k = st * (32 * ITEMS) + uint64_t(32 * i + gl);
          ok[i] = k < n;
          const uint64_t idx = ib + (ok[i] ? k : 0);  // clamped: ib < ie
          const int32_t v = items[idx];
          w[i] = ok[i] ? v : -2;
        }
        acc += probe.template lookup<ITEMS>(w, ok);
      }
      wave_sync();  // the tables are cleared for the next pair
      task = next;
      next = next2;
      ib = nib;
      ie = nie;
    }
  }
  flush_hits(acc, result);
}

__global__ void tc_reset(unsigned *queues, int nq,
                         unsigned long long *result) {
  const int i = threadIdx.x;
  if (i < nq) queues[i] = 0;
  if (i == 0) *result = 0;
}

constexpr int kWaves = 4;
constexpr int kSmallLogp = 8;    // 256 slots: d+ <= 128
constexpr int kMediumLogp = 11;  // 2048 slots: d+ <= 1024
constexpr int kBlockWaves = 8;

// Persistent grid: what fits the device at once, capped by `cap` (if set)
// and by the workgroups that can take a queue entry at all (`units`).
int grid_for(const void *kernel, int threads, size_t dyn_lds, int cus,
             int cap, int units) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads,
                                                   dyn_lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int g = per_cu * cus;
  if (cap > 0 && g > cap) g = cap;
  if (units > 0 && g > units) g = units;
  return g;
}

int waves_to_blocks(int groups, int waves) { return (groups + waves - 1) / waves; }

template <int ITEMS, bool STREAM>
int launch_all(const TcDevicePlan &p, const TcLaunch &l, hipStream_t stream) {
  const int merge_max =
      l.merge_max < kMergeMaxTlen ? l.merge_max : kMergeMaxTlen;
  // Heaviest classes first so their long tasks start early.
  if (p.ngroups[kTcBlock] > 0) {
    int logp = 6;
    while ((1 << logp) < 2 * p.class_max_tlen[kTcBlock]) ++logp;
    if (logp > kTcBlockLogpMax) return -1;
    const size_t lds = sizeof(int32_t) << logp;
    auto *kern = &tc_block_tables<kBlockWaves, ITEMS, STREAM>;
    const int grid = grid_for(reinterpret_cast<const void *>(kern),
                              kBlockWaves * 64, lds, l.cus, l.grid_cap,
                              p.ngroups[kTcBlock]);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlockWaves * 64), lds, stream,
                       p.col, p.ncol, p.tasks[kTcBlock], p.groups[kTcBlock],
                       p.ngroups[kTcBlock], p.partners, p.items, p.item_off,
                       p.queues + kTcBlock,
                       p.result, logp, l.cursor_min);
  }
  if (p.ngroups[kTcMedium] > 0) {
    auto *kern = &tc_wave_tables<kMediumLogp, kWaves, ITEMS, STREAM>;
    const int grid = grid_for(reinterpret_cast<const void *>(kern),
                              kWaves * 64, 0, l.cus, l.grid_cap,
                              waves_to_blocks(p.ngroups[kTcMedium], kWaves));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWaves * 64), 0, stream, p.col,
                       p.ncol, p.tasks[kTcMedium], p.groups[kTcMedium],
                       p.ngroups[kTcMedium], p.partners, p.items,
                       p.item_off, p.queues + kTcMedium, p.result,
                       l.cursor_min, merge_max);
  }
  if (STREAM && l.half_small && merge_max == 0 && p.ngroups[kTcSmall] > 0 &&
      p.class_max_tlen[kTcSmall] <= kTcSmallMaxDeg) {
    auto *kern = &tc_half_tables<kWaves, ITEMS>;
    const int grid = grid_for(reinterpret_cast<const void *>(kern),
                              kWaves * 64, 0, l.cus, l.grid_cap,
                              waves_to_blocks(p.ngroups[kTcSmall], kWaves));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWaves * 64), 0, stream, p.col,
                       p.tasks[kTcSmall], p.groups[kTcSmall],
                       p.ngroups[kTcSmall], p.items, p.item_off,
                       p.queues + kTcSmall, p.result);
  } else if (p.ngroups[kTcSmall] > 0) {
    auto *kern = &tc_wave_tables<kSmallLogp, kWaves, ITEMS, STREAM>;
    const int grid = grid_for(reinterpret_cast<const void *>(kern),
                              kWaves * 64, 0, l.cus, l.grid_cap,
                              waves_to_blocks(p.ngroups[kTcSmall], kWaves));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWaves * 64), 0, stream, p.col,
                       p.ncol, p.tasks[kTcSmall], p.groups[kTcSmall],
                       p.ngroups[kTcSmall], p.partners, p.items, p.item_off,
                       p.queues + kTcSmall, p.result, l.cursor_min,
                       merge_max);
  }
  if (p.ngroups[kTcGlobal] > 0) {
    auto *kern = &tc_global_search<kWaves, ITEMS, STREAM>;
    const int grid = grid_for(reinterpret_cast<const void *>(kern),
                              kWaves * 64, 0, l.cus, l.grid_cap,
                              waves_to_blocks(p.ngroups[kTcGlobal], kWaves));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWaves * 64), 0, stream, p.col,
                       p.ncol, p.tasks[kTcGlobal], p.groups[kTcGlobal],
                       p.ngroups[kTcGlobal], p.partners, p.items, p.item_off,
                       p.queues + kTcGlobal, p.result, l.cursor_min);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void partner_lens(const TcPartner *__restrict__ partners,
                             uint64_t np, uint64_t *__restrict__ lens) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i <= np;
       i += uint64_t(gridDim.x) * blockDim.x)
    lens[i] = i < np ? partners[i].len : 0u;
}

// One wave per batch of 64 partners: item g of the batch (its lists
// concatenated) belongs to the first partner o whose inclusive length sum
// exceeds g, and goes to items[item_off[first] + g].
constexpr int kLayoutWaves = 4;
__global__ __launch_bounds__(kLayoutWaves * 64) void layout_lists(
    const int32_t *__restrict__ col, const TcPartner *__restrict__ partners,
    uint64_t np, const uint64_t *__restrict__ item_off,
    int32_t *__restrict__ items) {
  __shared__ int32_t s_incl[kLayoutWaves][64];
  __shared__ uint32_t s_beg[kLayoutWaves][64];
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  const uint64_t nb = (np + 63) / 64;
  for (uint64_t b = uint64_t(blockIdx.x) * kLayoutWaves + wv; b < nb;
       b += uint64_t(gridDim.x) * kLayoutWaves) {
    const uint64_t i = b * 64 + lane;
    uint32_t beg = 0;
    int len = 0;
    if (i < np) {
      const TcPartner pr = partners[i];
      beg = pr.beg;
      len = static_cast<int>(pr.len);
    }
    const int incl = wave_incl_scan(len);
    const int total = __builtin_amdgcn_readlane(incl, 63);
    s_incl[wv][lane] = incl;
    s_beg[wv][lane] = beg;
    wave_sync();
    int32_t *out = items + item_off[b * 64];
    for (int g = lane; g < total; g += 64) {
      int lo = 0, hi = 63;  // incl[63] = total > g
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_incl[wv][mid] > g) hi = mid;
        else lo = mid + 1;
      }
      const int ex = lo > 0 ? s_incl[wv][lo - 1] : 0;
      out[g] = col[s_beg[wv][lo] + static_cast<uint32_t>(g - ex)];
    }
    wave_sync();  // before the next batch overwrites the sums
  }
}

__global__ void zero_words(uint32_t *p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

}  // namespace

int device_zero(void *p, size_t words, hipStream_t stream) {
  if (words == 0) return 0;
  const size_t blocks = std::min<size_t>((words + 255) / 256, 1024);
  hipLaunchKernelGGL(zero_words, dim3(static_cast<unsigned>(blocks)), dim3(256),
                     0, stream, static_cast<uint32_t *>(p), words);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int tc_device_init() {
  const int lds = static_cast<int>(sizeof(int32_t) << kTcBlockLogpMax);
  const void *kerns[] = {
      reinterpret_cast<const void *>(&tc_block_tables<kBlockWaves, 4, false>),
      reinterpret_cast<const void *>(&tc_block_tables<kBlockWaves, 8, false>),
      reinterpret_cast<const void *>(&tc_block_tables<kBlockWaves, 4, true>),
      reinterpret_cast<const void *>(&tc_block_tables<kBlockWaves, 8, true>)};
  for (const void *k : kerns)
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lds) != hipSuccess)
      return -1;
  return 0;
}

size_t tc_layout_scan_bytes(uint64_t np) {
  size_t b = 0;
  if (hipcub::DeviceScan::ExclusiveSum(
          nullptr, b, static_cast<const uint64_t *>(nullptr),
          static_cast<uint64_t *>(nullptr), np + 1) != hipSuccess)
    return 0;
  return b;
}

int tc_layout(const int32_t *col, const TcPartner *partners, uint64_t np,
              uint64_t *lens, uint64_t *item_off, int32_t *items,
              void *scan_tmp, size_t scan_bytes, int cus, hipStream_t stream) {
  const uint64_t blocks_l = std::min<uint64_t>((np + 256) / 256, 4096);
  hipLaunchKernelGGL(partner_lens, dim3(static_cast<unsigned>(blocks_l)),
                     dim3(256), 0, stream, partners, np, lens);
  size_t tb = scan_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, lens, item_off, np + 1,
                                       stream) != hipSuccess)
    return -1;
  if (np > 0) {
    const uint64_t batches = (np + 63) / 64;
    const uint64_t blocks = std::min<uint64_t>(
        (batches + kLayoutWaves - 1) / kLayoutWaves,
        static_cast<uint64_t>(cus) * 8);
    hipLaunchKernelGGL(layout_lists, dim3(static_cast<unsigned>(blocks)),
                       dim3(kLayoutWaves * 64), 0, stream, col, partners, np,
                       item_off, items);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int tc_device_launch(const TcDevicePlan &p, const TcLaunch &l,
                     hipStream_t stream) {
  if (l.reset)
    hipLaunchKernelGGL(tc_reset, dim3(1), dim3(64), 0, stream, p.queues,
                       kTcNumClasses, p.result);
  if (p.items && p.item_off)
    return l.items == 8 ? launch_all<8, true>(p, l, stream)
                        : launch_all<4, true>(p, l, stream);
  return l.items == 8 ? launch_all<8, false>(p, l, stream)
                      : launch_all<4, false>(p, l, stream);
}

}  // namespace mgpm

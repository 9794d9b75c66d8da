// Triangle-count device functions of the counting kernels (tc_kernels.hip,
// which describes the design).
#pragma once

#include <hip/hip_runtime.h>

#include "tc_device.hpp"
#include "wave.hpp"

namespace mgpm {
namespace tcdev {

constexpr int kMaxItems = 8;  // list items per lane per step (4 or 8)

// Streams one batch of up to 64 partner lists through `probe`: this lane's
// partner list starts at col[beg] and holds `len` items (0 for lanes past
// the batch).  The lists are concatenated and walked 64*ITEMS items per
// step.  `mark` is 64*ITEMS ints of wave-private LDS.  A batch whose lists
// average `cursor_min` items or more (never if <= 0) takes the owner-cursor
// path.  Returns this lane's hit count.
template <int ITEMS, class Probe>
__device__ __forceinline__ uint32_t stream_batch(
    const int32_t *__restrict__ col, uint32_t ncol, uint32_t beg, int len,
    lds_i32 *mark, const Probe &probe, int cursor_min, int lane) {
  // The start marks are written by one lane and read by another.  They are
  // LDS accesses (ds_*), which a wave executes in order; lds_order() keeps
  // the compiler from reordering or forwarding them across each handoff.
  uint32_t hits = 0;
  const int incl = wave_incl_scan(len);
  const int excl = incl - len;
  const int total = __builtin_amdgcn_readlane(incl, 63);
  if (cursor_min > 0 && total >= cursor_min * 64) {
    // Long lists on average: each lane walks its own owner cursor over
    // the batch's prefix sums (kept in LDS: a cursor advances in
    // divergent code, where cross-lane shuffles would read inactive
    // lanes).  Items lane + 64 i advance the cursor by at most a few
    // partners per step.
    lds_order();  // previous reads of the marks are done
    mark[lane] = incl;
    mark[64 + lane] = static_cast<int32_t>(beg);
    lds_order();
    int o = 0, oexcl = 0, onext = mark[0];
    uint32_t obeg = static_cast<uint32_t>(mark[64]);
    for (int k = 0; k < total; k += 64 * ITEMS) {
      int32_t w[ITEMS];
      bool ok[ITEMS];
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const int item = k + 64 * i + lane;
        ok[i] = item < total;
        if (ok[i]) {
          while (item >= onext) {
            ++o;
            oexcl = onext;
            onext = mark[o];
            obeg = static_cast<uint32_t>(mark[64 + o]);
          }
        }
        const uint32_t idx = obeg + static_cast<uint32_t>(item - oexcl);
        ok[i] = ok[i] && idx < ncol;
        w[i] = ok[i] ? col[idx] : -2;
      }
      hits += probe.template lookup<ITEMS>(w, ok);
    }
    return hits;
  }
  for (int k = 0; k < total; k += (64 * ITEMS)) {
    lds_order();  // previous reads of the marks are done
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) mark[lane + 64 * i] = 0;
    lds_order();
    if (excl > k && excl < k + (64 * ITEMS)) mark[excl - k] = 1;
    lds_order();
    unsigned long long starts[ITEMS];
    int own[ITEMS];  // a partner starts at this lane's item i
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const bool st = mark[lane + 64 * i] != 0;
      starts[i] = __ballot(st);
      own[i] = st ? 1 : 0;
    }
    // The partner owning item g is the last one with excl <= g; item g
    // of the batch is col[g + (beg - excl)] of its owner.  Every partner
    // list lies inside the DAG column array (checked when the plan is
    // built), so only items past the batch's end need a guard.
    int base = __popcll(__ballot(excl <= k)) - 1;
    const int delta = static_cast<int>(beg - static_cast<uint32_t>(excl));
    int32_t w[ITEMS];
    bool ok[ITEMS];
    if (k + 64 * ITEMS <= total) {  // a full step: no guards
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const int owner = base + bits_below(starts[i]) + own[i];
        base += __popcll(starts[i]);
        const uint32_t odelta = static_cast<uint32_t>(
            __builtin_amdgcn_ds_bpermute(owner << 2, delta));
        w[i] = col[static_cast<uint32_t>(k + 64 * i + lane) + odelta];
        ok[i] = true;
      }
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const int owner = base + bits_below(starts[i]) + own[i];
        base += __popcll(starts[i]);
        const uint32_t odelta = static_cast<uint32_t>(
            __builtin_amdgcn_ds_bpermute(owner << 2, delta));
        const int item = k + 64 * i + lane;
        ok[i] = item < total;
        // Unconditional gather (idle items read col[0]): no branch.
        const int32_t v =
            col[ok[i] ? static_cast<uint32_t>(item) + odelta : 0u];
        w[i] = ok[i] ? v : -2;
      }
    }
    hits += probe.template lookup<ITEMS>(w, ok);
  }
  return hits;
}

// Streams the partner lists [pbeg, pbeg+pcnt) of one task through `probe`,
// 64 partners at a time (batches first_batch, first_batch + batch_stride,
// ...).  `pre` (if non-null) is this lane's record of the first batch
// (valid for lanes below pcnt), already loaded by the caller; that batch
// is streamed outside the loop, so no join waits on loads issued after
// `pre`.  Returns this lane's hit count.
template <int ITEMS, class Probe>
__device__ __forceinline__ uint32_t stream_partners(
    const int32_t *__restrict__ col, uint32_t ncol,
    const TcPartner *__restrict__ partners, uint32_t pbeg, int pcnt,
    int first_batch, int batch_stride, int32_t *mark_lds,
    const Probe &probe, int cursor_min, const TcPartner *pre = nullptr) {
  lds_i32 *mark = as_lds(mark_lds);
  const int lane = lane64();
  uint32_t hits = 0;
  int pb = first_batch * 64;
  if (pre && pb < pcnt) {
    const bool mine = pb + lane < pcnt;
    hits += stream_batch<ITEMS>(col, ncol, mine ? pre->beg : 0u,
                                mine ? static_cast<int>(pre->len) : 0, mark,
                                probe, cursor_min, lane);
    pb += batch_stride * 64;
  }
  for (; pb < pcnt; pb += batch_stride * 64) {
    const int pi = pb + lane;
    uint32_t beg = 0;
    int len = 0;
    if (pi < pcnt) {
      const TcPartner pr = partners[pbeg + pi];
      beg = pr.beg;
      len = static_cast<int>(pr.len);
    }
    hits += stream_batch<ITEMS>(col, ncol, beg, len, mark, probe, cursor_min,
                                lane);
  }
  return hits;
}

// Streams the laid-out items [ib, ie) of one task (its partner lists,
// concatenated in partner order) through `probe`, 64 * ITEMS items per step;
// wave w of nw takes steps w, w + nw, ...  ib and ie are wave-uniform, so
// each load is one scalar base plus a lane offset.  Returns this lane's
// hit count.
template <int ITEMS, class Probe>
__device__ __forceinline__ uint32_t stream_items(
    const int32_t *__restrict__ items, uint64_t ib, uint64_t ie,
    const Probe &probe, int w, int nw) {
  const int lane = lane64();
  constexpr uint64_t kStep = 64ull * ITEMS;
  uint32_t hits = 0;
  for (uint64_t k = ib + uint64_t(w) * kStep; k < ie; k += uint64_t(nw) * kStep) {
    const int32_t *base = items + k;
    int32_t v[ITEMS];
    bool ok[ITEMS];
    if (k + kStep <= ie) {  // a full step: no guards
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        v[i] = base[64 * i + lane];
        ok[i] = true;
      }
    } else {
      const int left = static_cast<int>(ie - k);
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        ok[i] = 64 * i + lane < left;
        v[i] = ok[i] ? base[64 * i + lane] : -2;
      }
    }
    hits += probe.template lookup<ITEMS>(v, ok);
  }
  return hits;
}

// Open-addressing hash set in LDS, empty = -1.  Keys are vertex ids (>= 0),
// so an idle item (w = -2) never matches and its probe ends at the first
// empty slot (tables are at most half full).
//
// JOINT: all items of a lane probe together -- each round advances every
// item still probing and issues all their reads at once (finished items
// re-read their own slot), so a round costs one LDS round trip however
// many items miss, and the rounds run as long as the longest probe among
// the wave's items.  To keep that maximum short the sequence is quadratic
// (triangular offsets h, h+1, h+3, h+6, ..., which visit every slot of a
// power-of-two table), breaking up the clusters of linear probing; the
// table must be filled with the same sequence (table_insert, quad).
// Otherwise (tiny tables, where a second probe is rare) each item runs its
// own linear probe loop.
template <bool JOINT>
struct HashProbe {
  const int32_t *table;
  int logp;
  uint32_t mask;
  template <int ITEMS>
  __device__ __forceinline__ uint32_t lookup(const int32_t (&w)[ITEMS],
                                             const bool (&ok)[ITEMS]) const {
    uint32_t h[ITEMS];
    int32_t s[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      h[i] = slot_of(w[i], logp);
      s[i] = table[h[i]];
    }
    uint32_t hits = 0;
    if constexpr (JOINT) {
      uint32_t step = 1;  // triangular probe sequence h, h+1, h+3, h+6, ...
      for (;;) {
        bool more = false;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
          const bool miss = s[i] != w[i] && s[i] != -1;
          h[i] = miss ? ((h[i] + step) & mask) : h[i];
          more |= miss;
        }
        ++step;
        if (!more) break;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) s[i] = table[h[i]];
      }
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        hits += (ok[i] && s[i] == w[i]) ? 1u : 0u;
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (!ok[i]) continue;
        while (s[i] != w[i] && s[i] != -1) {
          h[i] = (h[i] + 1) & mask;
          s[i] = table[h[i]];
        }
        hits += s[i] == w[i] ? 1u : 0u;
      }
    }
    return hits;
  }
};

// Binary search in a sorted global list (tables too large for LDS).
struct SortedProbe {
  const int32_t *list;
  int len;
  template <int ITEMS>
  __device__ __forceinline__ uint32_t lookup(const int32_t (&w)[ITEMS],
                                             const bool (&ok)[ITEMS]) const {
    uint32_t hits = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if (!ok[i]) continue;
      int lo = 0, hi = len;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < w[i]) lo = mid + 1;
        else hi = mid;
      }
      hits += (lo < len && list[lo] == w[i]) ? 1u : 0u;
    }
    return hits;
  }
};

// Builds the hash set of col[tbeg, tbeg+tlen) into table[0, 2^logp) with
// `nthreads` cooperating threads of index `tid`.  Caller clears first.
// `quad`: the triangular probe sequence of HashProbe<true>, else linear.
__device__ __forceinline__ void table_insert(int32_t *table, int logp,
                                             const int32_t *__restrict__ col,
                                             uint32_t tbeg, int tlen, int tid,
                                             int nthreads, bool quad) {
  const uint32_t mask = (1u << logp) - 1u;
  for (int i = tid; i < tlen; i += nthreads) {
    const int32_t key = col[tbeg + i];
    uint32_t h = slot_of(key, logp);
    uint32_t step = 1;
    while (atomicCAS(&table[h], -1, key) != -1) {
      h = (h + (quad ? step : 1u)) & mask;
      ++step;
    }
  }
}

__device__ __forceinline__ void flush_hits(unsigned long long acc,
                                           unsigned long long *result) {
  acc = wave_sum_u64(acc);
  if (lane64() == 0 && acc) atomicAdd(result, acc);
}


// What a wave loads for a task before building its table: this lane's
// partner record of the first batch and this lane's owner items (lane,
// lane + 64; used for owners of up to 128 items).  The loads are issued
// unconditionally at clamped (valid) addresses and their values are
// selected only where they are used, so loads issued one task ahead stay
// in flight across the previous task's work (a select right after a load
// would wait for it).
struct TaskPre {
  TcPartner first{0u, 0u};
  int32_t key0 = 0, key1 = 0;
  uint64_t ib = 0, ie = 0;  // laid-out items of the task (STREAM)
};

constexpr int kTaskPreMaxTlen = 128;

// Tasks have pcnt >= 1 and tlen >= 1 (owners without partners get none).
template <bool STREAM = false>
__device__ __forceinline__ TaskPre load_task_pre(
    const int32_t *__restrict__ col, const TcPartner *__restrict__ partners,
    const TcTask &t, int lane, const uint64_t *__restrict__ item_off = nullptr) {
  TaskPre p;
  if constexpr (STREAM) {
    p.ib = item_off[t.pbeg];
    p.ie = item_off[t.pbeg + static_cast<uint32_t>(t.pcnt)];
  } else {
    p.first = partners[t.pbeg + static_cast<uint32_t>(min(lane, t.pcnt - 1))];
  }
  p.key0 = col[t.tbeg + static_cast<uint32_t>(min(lane, t.tlen - 1))];
  p.key1 = col[t.tbeg + static_cast<uint32_t>(min(lane + 64, t.tlen - 1))];
  return p;
}

// table_insert of up to two keys per lane already in registers (valid
// flags v0, v1).
__device__ __forceinline__ void table_insert_keys(int32_t *table, int logp,
                                                  int32_t k0, bool v0,
                                                  int32_t k1, bool v1,
                                                  bool quad) {
  const uint32_t mask = (1u << logp) - 1u;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int32_t key = k == 0 ? k0 : k1;
    if (!(k == 0 ? v0 : v1)) continue;
    uint32_t h = slot_of(key, logp);
    uint32_t step = 1;
    while (atomicCAS(&table[h], -1, key) != -1) {
      h = (h + (quad ? step : 1u)) & mask;
      ++step;
    }
  }
}

// Counts one task: builds the owner's hash set in the wave-private table
// (2^logp_max slots, all lanes of this wave) and streams its partners.  A
// table that would not fit falls back to binary search in global memory.
template <int ITEMS>
// `pre` holds what load_task_pre loaded for this task.
__device__ __forceinline__ uint32_t wave_task(const int32_t *__restrict__ col,
                                              uint32_t ncol,
                                              const TcPartner *__restrict__ partners,
                                              const TcTask &task, int32_t *table,
                                              int logp_max, int32_t *mark,
                                              int cursor_min,
                                              const TaskPre &pre) {
  const int lane = lane64();
  if (2 * task.tlen > (1 << logp_max)) {
    SortedProbe probe{col + task.tbeg, task.tlen};
    return stream_partners<ITEMS>(col, ncol, partners, task.pbeg, task.pcnt, 0,
                                  1, mark, probe, cursor_min, &pre.first);
  }
  // 8 slots per key where the table fits (load <= 1/8): fewer probe rounds.
  int logp = ceil_log2(8 * task.tlen);
  logp = logp < 6 ? 6 : (logp > logp_max ? logp_max : logp);
  const int P = 1 << logp;
  for (int i = lane; i < P; i += 64) table[i] = -1;
  wave_sync();
  if (task.tlen <= kTaskPreMaxTlen)
    table_insert_keys(table, logp, pre.key0, lane < task.tlen, pre.key1,
                      lane + 64 < task.tlen, logp_max > 8);
  else
    table_insert(table, logp, col, task.tbeg, task.tlen, lane, 64,
                 logp_max > 8);
  wave_sync();
  // Tables of up to 256 slots (small owners) probe item by item.
  if (logp_max <= 8) {
    HashProbe<false> probe{table, logp, static_cast<uint32_t>(P - 1)};
    const uint32_t hits = stream_partners<ITEMS>(
        col, ncol, partners, task.pbeg, task.pcnt, 0, 1, mark, probe,
        cursor_min, &pre.first);
    wave_sync();
    return hits;
  }
  HashProbe<true> probe{table, logp, static_cast<uint32_t>(P - 1)};
  const uint32_t hits = stream_partners<ITEMS>(
      col, ncol, partners, task.pbeg, task.pcnt, 0, 1, mark, probe, cursor_min,
      &pre.first);
  wave_sync();
  return hits;
}

// Owners of up to kMergeMaxTlen items can skip the table: the owner list
// goes to LDS as it is (sorted), and each lane merges one partner list
// (sorted, at most as long as the owner's by the owner rule) against it
// with two cursors, the next partner item loaded one step ahead.  Lists
// come from the DAG (partners) or, when `items` is set, from the layout.
constexpr int kMergeMaxTlen = 64;

__device__ __forceinline__ uint32_t wave_task_merge(
    const int32_t *__restrict__ col, const TcPartner *__restrict__ partners,
    const int32_t *__restrict__ items, const uint64_t *__restrict__ item_off,
    const TcTask &task, int32_t *own_lds, const TaskPre &pre) {
  const int lane = lane64();
  lds_i32 *own = as_lds(own_lds);
  const int tlen = task.tlen;  // 1 .. kMergeMaxTlen
  if (lane < tlen) own[lane] = pre.key0;
  wave_sync();
  uint32_t hits = 0;
  for (int pb = 0; pb < task.pcnt; pb += 64) {
    const int pi = pb + lane;
    if (pi >= task.pcnt) continue;
    const int32_t *L;
    int len;
    const uint32_t p = task.pbeg + static_cast<uint32_t>(pi);
    if (items) {
      const uint64_t b0 = item_off[p];
      L = items + b0;
      len = static_cast<int>(item_off[p + 1] - b0);
    } else {
      const TcPartner pr = partners[p];
      L = col + pr.beg;
      len = static_cast<int>(pr.len);
    }
    if (len <= 0) continue;
    int i = 0, j = 0;
    int32_t a = own[0];
    int32_t b = L[0];
    int32_t bn = len > 1 ? L[1] : 0;
    for (;;) {
      const bool ai = a <= b, bj = b <= a;  // both on a hit
      hits += (ai && bj) ? 1u : 0u;
      if (ai && ++i >= tlen) break;
      if (bj && ++j >= len) break;
      if (ai) a = own[i];
      if (bj) {
        b = bn;
        bn = j + 1 < len ? L[j + 1] : 0;
      }
    }
  }
  wave_sync();  // the owner list is rewritten by the next task
  return hits;
}

// The first step of a task's laid-out items (64 * ITEMS from ib), loaded
// ahead of the task: every load is issued, at an address clamped into the
// task's items (a task has at least one), and the flags say which count.
template <int ITEMS>
struct FirstStep {
  int32_t v[ITEMS];
};

template <int ITEMS>
__device__ __forceinline__ FirstStep<ITEMS> load_first_step(
    const int32_t *__restrict__ items, uint64_t ib, uint64_t ie, int lane) {
  FirstStep<ITEMS> f;
  const uint64_t last = ie - 1;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t idx = ib + uint64_t(64 * i + lane);
    f.v[i] = items[idx < last ? idx : last];
  }
  return f;
}

// wave_task over the task's laid-out items (pre.ib .. pre.ie) instead of
// its partner lists; `first` (if non-null) holds its first step, loaded
// ahead by load_first_step.
template <int ITEMS>
__device__ __forceinline__ uint32_t wave_task_stream(
    const int32_t *__restrict__ col, const int32_t *__restrict__ items,
    const TcTask &task, int32_t *table, int logp_max, const TaskPre &pre,
    const FirstStep<ITEMS> *first = nullptr) {
  const int lane = lane64();
  uint64_t ib = wave_first_u64(pre.ib);
  const uint64_t ie = wave_first_u64(pre.ie);
  if (2 * task.tlen > (1 << logp_max)) {
    SortedProbe probe{col + task.tbeg, task.tlen};
    return stream_items<ITEMS>(items, ib, ie, probe, 0, 1);
  }
  int logp = ceil_log2(8 * task.tlen);
  logp = logp < 6 ? 6 : (logp > logp_max ? logp_max : logp);
  const int P = 1 << logp;
  for (int i = lane; i < P; i += 64) table[i] = -1;
  wave_sync();
  if (task.tlen <= kTaskPreMaxTlen)
    table_insert_keys(table, logp, pre.key0, lane < task.tlen, pre.key1,
                      lane + 64 < task.tlen, logp_max > 8);
  else
    table_insert(table, logp, col, task.tbeg, task.tlen, lane, 64,
                 logp_max > 8);
  wave_sync();
  uint32_t hits = 0;
  int32_t fv[ITEMS];
  bool fok[ITEMS];
  if (first) {  // the first step from registers, the rest streamed
    const int left = ie - ib < uint64_t(64 * ITEMS)
                         ? static_cast<int>(ie - ib) : 64 * ITEMS;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      fok[i] = 64 * i + lane < left;
      fv[i] = fok[i] ? first->v[i] : -2;
    }
  }
  if (logp_max <= 8) {
    HashProbe<false> probe{table, logp, static_cast<uint32_t>(P - 1)};
    if (first) {
      hits = probe.template lookup<ITEMS>(fv, fok);
      ib += 64 * ITEMS;
    }
    hits += stream_items<ITEMS>(items, ib, ie, probe, 0, 1);
  } else {
    HashProbe<true> probe{table, logp, static_cast<uint32_t>(P - 1)};
    if (first) {
      hits = probe.template lookup<ITEMS>(fv, fok);
      ib += 64 * ITEMS;
    }
    hits += stream_items<ITEMS>(items, ib, ie, probe, 0, 1);
  }
  wave_sync();
  return hits;
}

}  // namespace tcdev
}  // namespace mgpm

// Pair passes of the 5-vertex census: the subgraph counts behind P2, P3,
// P5 and P6 that no per-edge or per-triangle sum gives -- wheels W4 (a hub
// and a 4-cycle in its neighbourhood), K2,3 and K2,3 plus an edge between
// two of its three middle vertices.  All three are sums over vertex pairs
// {a, c} of functions of the common neighbourhood S = N(a) ∩ N(c):
//   K2,3 = sum C(|S|, 3),  K2,3+e = sum E(S) (|S| - 2),
//   W4 = 1/2 sum wedges(S) (each wheel has two rim diagonals).
// Every occurrence is charged to its lowest-ranked vertex in the DAG order
// (degree, then id; graph.cc:205-235), whose pattern neighbours all lie in
// its DAG out-neighbourhood N+ (at most ~sqrt(2m) vertices):
//   * anchor pass, per anchor a and every c above a: the mask
//     M = N+(a) ∩ N(c) over N+(a) and lo = |{z below a: z->a, z->c}|, so
//     codeg(a, c) = lo + |M| (common neighbours above a lie in N+(a));
//     K2,3 += C(codeg, 3); wheels whose lowest vertex is the rim vertex a
//     (opposite rim vertex c) += wedges of G[M]; diamonds whose lowest
//     vertex is the tip a (other tip c, hinge inside M) give
//     E(G[M]) (codeg - 2) of K2,3+e; and codeg of every pair with lo >= 1
//     goes into a table hashed per a;
//   * out-neighbourhood pass, per anchor x: wheels whose lowest vertex is
//     the hub x are 4-cycles inside G[N+(x)]; diamonds whose lowest vertex
//     is the hinge end x (tips p, q and hinge end y in N+(x)) give
//     |N+(x) ∩ N(p) ∩ N(q)| (codeg(p, q) - 2) of K2,3+e, codeg from the
//     table (x lies below p and q, so the pair is there).
// G[N+(a)] is held as adjacency bit rows, in LDS for d+ up to 512.  The
// passes run on the census graph, relabelled in DAG order (census_relabel):
// ids compare as ranks, and each sorted row lists the vertices below its
// own first, then N+.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>

#include "census.hpp"
#include "choose.hpp"
#include "wave.hpp"

namespace mgpm {

namespace {

constexpr int kT = 256;  // threads per workgroup
constexpr int kTW = kT / 64;

// b is ranked above a in DAG order: the census graph is relabelled in that
// order (census_relabel), so the larger id.
__device__ __forceinline__ bool above(int32_t a, int32_t b) { return b > a; }

template <class P>
__device__ __forceinline__ int lower_bound_p(P list, int len, int32_t key) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (list[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t tab_slot(int32_t key, int logc) {
  return logc == 0 ? 0
                   : (static_cast<uint64_t>(static_cast<uint32_t>(key)) *
                      0x9E3779B97F4A7C15ull) >> (64 - logc);
}

__device__ __forceinline__ int log2_pow2(uint64_t cap) {
  return 63 - __clzll(static_cast<long long>(cap));
}

// Entries: (key + 1) << 32 | value; 0 = empty.  Load factor <= 1/2 when
// the count pass and the anchor pass agree; false if no slot was free.
__device__ bool tab_insert(unsigned long long *tab, uint64_t cap, int32_t key,
                           uint32_t val) {
  if (cap == 0) return false;
  const unsigned long long e =
      (static_cast<unsigned long long>(static_cast<uint32_t>(key) + 1u) << 32) |
      val;
  const int lc = log2_pow2(cap);
  uint64_t s = tab_slot(key, lc);
  for (uint64_t probes = 0; probes < cap; ++probes) {
    if (atomicCAS(&tab[s], 0ull, e) == 0ull) return true;
    s = (s + 1) & (cap - 1);
  }
  return false;
}

__device__ bool tab_find(const unsigned long long *tab, uint64_t cap,
                         int32_t key, uint32_t *val) {
  if (cap == 0) return false;
  const unsigned long long k1 = static_cast<uint32_t>(key) + 1u;
  const int lc = log2_pow2(cap);
  uint64_t s = tab_slot(key, lc);
  for (uint64_t probes = 0; probes < cap; ++probes) {
    const unsigned long long e = tab[s];
    if (e == 0) return false;
    if ((e >> 32) == k1) {
      *val = static_cast<uint32_t>(e);
      return true;
    }
    s = (s + 1) & (cap - 1);
  }
  return false;
}

__device__ __forceinline__ unsigned long long wg_sum(unsigned long long x,
                                                     unsigned long long *red) {
  x = wave_sum_u64(x);
  __syncthreads();
  if (lane64() == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  unsigned long long t = 0;
  for (int k = 0; k < kTW; ++k) t += red[k];
  return t;
}

// t per undirected CSR arc: the DAG arc between its ends, found in the
// DAG row of the lower-ranked end.
__global__ void c5_t_csr(Census5Args c) {
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < c.nnz;
       j += int64_t(gridDim.x) * blockDim.x) {
    const int32_t v = c.src[j], u = c.col[j];
    const bool vu = above(v, u);
    const int32_t lo = vu ? v : u, hi = vu ? u : v;
    const int64_t b = c.drp[lo];
    const int len = static_cast<int>(c.drp[lo + 1] - b);
    const int k = lower_bound_p(c.dcol + b, len, hi);
    c.t_csr[j] = c.t[b + k];
  }
}

// Sum over vertices of C(t_v, 2), and of t_v s_v as a sum over CSR arcs
// (v, x) of t_v (d_x - 1).
__global__ void c5_flat(Census5Args c) {
  unsigned long long ts = 0, t2 = 0;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t j = i0; j < c.nnz; j += stride) {
    const unsigned long long tv = c.tv[c.src[j]] >> 1;
    ts += tv * static_cast<unsigned long long>(c.deg[c.col[j]] - 1);
  }
  for (int64_t v = i0; v < c.n; v += stride) t2 += choose2(c.tv[v] >> 1);
  ts = wave_sum_u64(ts);
  t2 = wave_sum_u64(t2);
  if (lane64() == 0) {
    if (ts) atomicAdd(&c.acc[kCensusTriS], ts);
    if (t2) atomicAdd(&c.acc[kCensusTv2], t2);
  }
}

__device__ __forceinline__ int class_of(int dp) {
  int k = 0;
  while (k < kC5Classes - 1 && dp > 64 * kC5Words[k]) ++k;
  return k;
}

// Class sizes and the scatter of the anchors into their classes, with one
// atomic per class per wave (ballots over the wave's 64 vertices).  The
// loops run whole waves to the end (inactive lanes hold class -1).
__global__ void c5_class_count(Census5Args c) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v0 = int64_t(blockIdx.x) * blockDim.x; v0 < c.n; v0 += stride) {
    const int64_t v = v0 + threadIdx.x;
    const int k = v < c.n ? class_of(static_cast<int>(c.drp[v + 1] - c.drp[v]))
                          : -1;
    for (int q = 0; q < kC5Classes; ++q) {
      const unsigned long long m = __ballot(k == q);
      if (lane64() == 0 && m)
        atomicAdd(&c.cls_count[q], static_cast<unsigned>(__popcll(m)));
    }
  }
}

__global__ void c5_class_scatter(Census5Args c) {
  unsigned start[kC5Classes];
  unsigned s = 0;
  for (int k = 0; k < kC5Classes; ++k) {
    start[k] = s;
    s += c.cls_count[k];
  }
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t v0 = int64_t(blockIdx.x) * blockDim.x; v0 < c.n; v0 += stride) {
    const int64_t v = v0 + threadIdx.x;
    const int k = v < c.n ? class_of(static_cast<int>(c.drp[v + 1] - c.drp[v]))
                          : -1;
    for (int q = 0; q < kC5Classes; ++q) {
      const unsigned long long m = __ballot(k == q);
      if (!m) continue;
      unsigned base = 0;
      if (lane64() == 0)
        base = atomicAdd(&c.cls_fill[q], static_cast<unsigned>(__popcll(m)));
      base = __shfl(base, 0, 64);
      if (k == q)
        c.anchors[start[q] + base + bits_below(m)] = static_cast<int32_t>(v);
    }
  }
}

// Table sizes: per anchor a, the distinct c above a with a common
// neighbour z below a (z->a, z->c); capacity a power of two >= 2x that.
__global__ __launch_bounds__(kT) void c5_tab_count(Census5Args c) {
  uint32_t *stamp = c.stamp + size_t(blockIdx.x) * size_t(c.n);
  __shared__ unsigned s_a, s_cnt;
  const int lane = lane64(), wv = threadIdx.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) {
      s_a = atomicAdd(&c.cursor[0], 1u);
      s_cnt = 0;
    }
    __syncthreads();
    const unsigned ua = s_a;
    if (ua >= static_cast<unsigned>(c.n)) break;
    const int32_t a = static_cast<int32_t>(ua);
    const uint32_t tag = ua + 1u;
    // z below a: the first d(a) - d+(a) entries of a's row in DAG order.
    const int64_t ea = c.rowptr[a + 1] - (c.drp[a + 1] - c.drp[a]);
    for (int64_t k = c.rowptr[a] + wv; k < ea; k += kTW) {
      const int32_t z = c.col[k];
      const int64_t bz = c.drp[z];
      const int dz = static_cast<int>(c.drp[z + 1] - bz);
      for (int q = lane; q < dz; q += 64) {
        const int32_t cc = c.dcol[bz + q];
        if (!above(a, cc)) continue;
        if (atomicExch(&stamp[cc], tag) != tag) atomicAdd(&s_cnt, 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long n2 = 2ull * s_cnt;
      unsigned long long cap = 0;
      if (n2) {
        cap = 1;
        while (cap < n2) cap <<= 1;
      }
      c.tab_cap[a] = cap;
    }
    __syncthreads();  // s_a, s_cnt reused
  }
}

// Adjacency bit rows of G[N+(a)]: bit j of row i iff nb[j] ~ nb[i].
// `rows` (stride rs words) must be zero; built with atomics.
template <class NB, class ROWS>
__device__ void build_rows(const Census5Args &c, NB nb, int dp, ROWS rows,
                           int rs) {
  const int lane = lane64(), wv = threadIdx.x >> 6;
  for (int i = wv; i < dp; i += kTW) {
    const int32_t u = nb[i];
    const int64_t bu = c.drp[u];
    const int du = static_cast<int>(c.drp[u + 1] - bu);
    for (int k = lane; k < du; k += 64) {
      const int32_t w = c.dcol[bu + k];
      const int j = lower_bound_p(nb, dp, w);
      if (j < dp && nb[j] == w) {
        atomicOr(&rows[int64_t(i) * rs + (j >> 6)], 1ull << (j & 63));
        atomicOr(&rows[int64_t(j) * rs + (i >> 6)], 1ull << (i & 63));
      }
    }
  }
}

// Reads a row word built by atomics: LDS words directly, global words past
// the vector L1 (the atomics ran in L2).
__device__ __forceinline__ unsigned long long row_word(
    const unsigned long long *r, bool global) {
  return global ? __hip_atomic_load(r, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT)
                : *r;
}

// W > 0: anchors with d+ <= 64 W, rows in LDS, masks of W words.
// W == 0: the last class, rows in the workgroup's global slab, masks of
// gen_words words.
template <int W>
__global__ __launch_bounds__(kT) void c5_anchor(Census5Args c, int cls) {
  constexpr int LW = W > 0 ? W : 1;
  __shared__ int32_t s_nb[W > 0 ? 64 * W : 1];
  __shared__ unsigned long long s_rows[W > 0 ? 64 * W * W : 1];
  __shared__ unsigned s_a, s_nt;
  __shared__ unsigned long long s_red[kTW];
  const int lane = lane64(), wv = threadIdx.x >> 6;
  const int rs = W > 0 ? W : c.gen_words;
  const size_t n = static_cast<size_t>(c.n);
  unsigned long long *mask = c.mask + size_t(blockIdx.x) * n * rs;
  uint32_t *lo = c.lo + size_t(blockIdx.x) * n;
  uint32_t *stamp = c.stamp + size_t(blockIdx.x) * n;
  int32_t *list = c.list + size_t(blockIdx.x) * n;
  unsigned long long *grows = c.grows + size_t(blockIdx.x) * c.grows_len;
  unsigned cls_begin = 0;
  for (int k = 0; k < cls; ++k) cls_begin += c.cls_count[k];
  const unsigned cls_n = c.cls_count[cls];
  unsigned long long k23 = 0, w4 = 0, a16 = 0, miss = 0;
  (void)LW;
  for (;;) {
    if (threadIdx.x == 0) {
      s_a = atomicAdd(&c.cursor[1 + cls], 1u);
      s_nt = 0;
    }
    __syncthreads();
    const unsigned ai = s_a;
    if (ai >= cls_n) break;
    const int32_t a = c.anchors[cls_begin + ai];
    const uint32_t tag = static_cast<uint32_t>(c.n) + 1u + static_cast<uint32_t>(a);
    const int64_t ba = c.drp[a];
    const int dp = static_cast<int>(c.drp[a + 1] - ba);
    const int ru = (dp + 63) >> 6;  // words in use
    const int32_t *gnb = c.dcol + ba;
    unsigned long long *rows = W > 0 ? s_rows : grows;
    if (W > 0)
      for (int i = threadIdx.x; i < dp; i += kT) s_nb[i] = gnb[i];
    for (int64_t i = threadIdx.x; i < int64_t(dp) * rs; i += kT) rows[i] = 0;
    __syncthreads();
    if (W > 0) build_rows(c, s_nb, dp, rows, rs);
    else build_rows(c, gnb, dp, rows, rs);
    // Masks: bit i of M[c] for every c above a adjacent to nb[i].
    for (int i = wv; i < dp; i += kTW) {
      const int32_t y = W > 0 ? s_nb[i] : gnb[i];
      const unsigned long long bit = 1ull << (i & 63);
      const int wd = i >> 6;
      // c above a: the tail of y's sorted row past a (a itself lies in the
      // row, y being above a).
      const int64_t ey = c.rowptr[y + 1];
      const int64_t ky = wave_first_greater(c.col, a, c.rowptr[y], ey);
      for (int64_t k = ky + lane; k < ey; k += 64) {
        const int32_t cc = c.col[k];
        const unsigned long long old =
            atomicOr(&mask[size_t(cc) * rs + wd], bit);
        // One-word masks: the first bit set on c (its mask and count were
        // zero) lists it; wider masks stamp c with the anchor.
        if (W == 1 ? old == 0ull : atomicExch(&stamp[cc], tag) != tag)
          list[atomicAdd(&s_nt, 1u)] = cc;
      }
    }
    if (W == 1) __syncthreads();  // every mask final before the counts
    // lo: common neighbours below a, through every z below a (the first
    // d(a) - d+(a) entries of a's row in DAG order).
    const int64_t ea = c.rowptr[a + 1] - dp;
    for (int64_t k = c.rowptr[a] + wv; k < ea; k += kTW) {
      const int32_t z = c.col[k];
      const int64_t bz = c.drp[z];
      const int dz = static_cast<int>(c.drp[z + 1] - bz);
      for (int q = lane; q < dz; q += 64) {
        const int32_t cc = c.dcol[bz + q];
        if (!above(a, cc)) continue;
        const uint32_t old = atomicAdd(&lo[cc], 1u);
        bool first;
        if (W == 1)  // listed already if its mask has a bit
          first = old == 0u &&
                  __hip_atomic_load(&mask[size_t(cc)], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == 0ull;
        else
          first = atomicExch(&stamp[cc], tag) != tag;
        if (first) list[atomicAdd(&s_nt, 1u)] = cc;
      }
    }
    __syncthreads();
    const unsigned nt = s_nt;
    const uint64_t toff = c.tab_off[a];
    const uint64_t tcap = c.tab_off[a + 1] - toff;
    for (unsigned k = threadIdx.x; k < nt; k += kT) {
      const int32_t cc = list[k];
      unsigned long long *mc = mask + size_t(cc) * rs;
      // Read and clear in one atomic each (made by atomics, left zeroed).
      const uint32_t l = atomicExch(&lo[cc], 0u);
      unsigned long long m[W > 0 ? W : 1];
      unsigned pc = 0;
      if (W > 0) {
#pragma unroll
        for (int w = 0; w < LW; ++w) {  // words past ru are zero
          m[w] = atomicExch(&mc[w], 0ull);
          pc += __popcll(m[w]);
        }
      } else {
        for (int w = 0; w < ru; ++w)
          pc += __popcll(__hip_atomic_load(&mc[w], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
      }
      const unsigned long long cd = static_cast<unsigned long long>(l) + pc;
      k23 += choose3(cd);
      if (pc >= 2) {
        unsigned long long e2 = 0, wg = 0;
        if (W > 0) {
#pragma unroll
          for (int w = 0; w < LW; ++w) {
            for (unsigned long long b = m[w]; b; b &= b - 1) {
              const int p = w * 64 + __builtin_ctzll(b);
              unsigned dm = 0;
#pragma unroll
              for (int w2 = 0; w2 < LW; ++w2)
                dm += __popcll(rows[p * LW + w2] & m[w2]);
              e2 += dm;
              wg += choose2(dm);
            }
          }
        } else {
          for (int w = 0; w < ru; ++w) {
            const unsigned long long mw = __hip_atomic_load(
                &mc[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned long long b = mw; b; b &= b - 1) {
              const int p = w * 64 + __builtin_ctzll(b);
              unsigned dm = 0;
              for (int w2 = 0; w2 < ru; ++w2) {
                const unsigned long long m2 = __hip_atomic_load(
                    &mc[w2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                dm += __popcll(row_word(&rows[int64_t(p) * rs + w2], true) &
                               m2);
              }
              e2 += dm;
              wg += choose2(dm);
            }
          }
        }
        w4 += wg;
        a16 += (e2 >> 1) * (cd - 2);
      }
      if (W == 0)
        for (int w = 0; w < ru; ++w) atomicExch(&mc[w], 0ull);
      if (l && !tab_insert(c.table + toff, tcap, cc, static_cast<uint32_t>(cd)))
        ++miss;
    }
    __syncthreads();  // s_a, s_nt, the list and the rows are reused
  }
  k23 = wave_sum_u64(k23);
  w4 = wave_sum_u64(w4);
  a16 = wave_sum_u64(a16);
  miss = wave_sum_u64(miss);
  if (lane == 0) {
    if (k23) atomicAdd(&c.acc[kCensusK23], k23);
    if (w4) atomicAdd(&c.acc[kCensusW4], w4);
    if (a16) atomicAdd(&c.acc[kCensusK23e], a16);
    if (miss) atomicAdd(&c.acc[kCensusTableMiss], miss);
  }
  (void)s_red;
}

template <int W>
__global__ __launch_bounds__(kT) void c5_outnb(Census5Args c, int cls) {
  __shared__ int32_t s_nb[W > 0 ? 64 * W : 1];
  __shared__ unsigned long long s_rows[W > 0 ? 64 * W * W : 1];
  __shared__ unsigned s_a;
  __shared__ unsigned long long s_red[kTW];
  const int lane = lane64();
  const int rs = W > 0 ? W : c.gen_words;
  unsigned long long *grows = c.grows + size_t(blockIdx.x) * c.grows_len;
  unsigned cls_begin = 0;
  for (int k = 0; k < cls; ++k) cls_begin += c.cls_count[k];
  const unsigned cls_n = c.cls_count[cls];
  unsigned long long w4 = 0, b16 = 0, miss = 0;
  for (;;) {
    if (threadIdx.x == 0) s_a = atomicAdd(&c.cursor[1 + kC5Classes + cls], 1u);
    __syncthreads();
    const unsigned ai = s_a;
    if (ai >= cls_n) break;
    const int32_t x = c.anchors[cls_begin + ai];
    const int64_t bx = c.drp[x];
    const int dp = static_cast<int>(c.drp[x + 1] - bx);
    const int ru = (dp + 63) >> 6;
    const int32_t *gnb = c.dcol + bx;
    unsigned long long *rows = W > 0 ? s_rows : grows;
    unsigned long long cyc = 0;
    if (dp >= 2) {
      if (W > 0)
        for (int i = threadIdx.x; i < dp; i += kT) s_nb[i] = gnb[i];
      for (int64_t i = threadIdx.x; i < int64_t(dp) * rs; i += kT) rows[i] = 0;
      __syncthreads();
      if (W > 0) build_rows(c, s_nb, dp, rows, rs);
      else build_rows(c, gnb, dp, rows, rs);
      __syncthreads();
      const int64_t np = int64_t(dp) * dp;
      for (int64_t k = threadIdx.x; k < np; k += kT) {
        const int i = static_cast<int>(k / dp), j = static_cast<int>(k - int64_t(i) * dp);
        if (j <= i) continue;
        unsigned h = 0;
        for (int w = 0; w < ru; ++w)
          h += __popcll(row_word(&rows[int64_t(i) * rs + w], W == 0) &
                        row_word(&rows[int64_t(j) * rs + w], W == 0));
        if (!h) continue;
        cyc += choose2(h);
        const int32_t p = W > 0 ? s_nb[i] : gnb[i];
        const int32_t q = W > 0 ? s_nb[j] : gnb[j];
        const bool pq = above(p, q);
        const int32_t lo_v = pq ? p : q, hi_v = pq ? q : p;
        const uint64_t off = c.tab_off[lo_v];
        uint32_t cd = 0;
        if (tab_find(c.table + off, c.tab_off[lo_v + 1] - off, hi_v, &cd))
          b16 += static_cast<unsigned long long>(h) * (cd - 2ull);
        else
          ++miss;
      }
    }
    // Each 4-cycle of G[N+(x)] was seen at both of its diagonals.
    const unsigned long long tot = wg_sum(cyc, s_red);
    if (threadIdx.x == 0) w4 += tot >> 1;
    __syncthreads();  // s_a and the rows are reused
  }
  b16 = wave_sum_u64(b16);
  miss = wave_sum_u64(miss);
  if (lane == 0) {
    if (b16) atomicAdd(&c.acc[kCensusK23e], b16);
    if (miss) atomicAdd(&c.acc[kCensusTableMiss], miss);
  }
  if (threadIdx.x == 0 && w4) atomicAdd(&c.acc[kCensusW4], w4);
}

// Relabelling in DAG order (census_relabel): degrees and ids, then ranks.
__global__ void c5_degrees(const int64_t *__restrict__ rowptr, int32_t n,
                           uint32_t *__restrict__ deg,
                           uint32_t *__restrict__ ids) {
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
       v += int64_t(gridDim.x) * blockDim.x) {
    deg[v] = static_cast<uint32_t>(rowptr[v + 1] - rowptr[v]);
    ids[v] = static_cast<uint32_t>(v);
  }
}

// rank[perm[i]] = i: perm lists the vertices in DAG order; lens[i] = the
// degree of rank i (lens[n] = 0, for the scan into row pointers).
__global__ void c5_ranks(const uint32_t *__restrict__ perm,
                         const uint32_t *__restrict__ dsort, int32_t n,
                         uint32_t *__restrict__ rank,
                         int64_t *__restrict__ lens) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i <= n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (i < n) rank[perm[i]] = static_cast<uint32_t>(i);
    lens[i] = i < n ? static_cast<int64_t>(dsort[i]) : 0;
  }
}

// One key per arc, one wave per row: the rank of its row above the rank
// of its head (lg bits each).
__global__ void c5_arc_keys(const int64_t *__restrict__ rowptr,
                            const int32_t *__restrict__ col, int32_t n,
                            const uint32_t *__restrict__ rank, int lg,
                            unsigned long long *__restrict__ keys) {
  const int lane = lane64();
  const int64_t nwaves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  for (int64_t v = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
       v < n; v += nwaves) {
    const unsigned long long r = static_cast<unsigned long long>(rank[v])
                                 << lg;
    const int64_t e = rowptr[v + 1];
    for (int64_t i = rowptr[v] + lane; i < e; i += 64)
      keys[i] = r | rank[col[i]];
  }
}

// The heads' new ids (their ranks), in the sorted order.
__global__ void c5_arc_heads(const unsigned long long *__restrict__ keys,
                             int64_t nnz, unsigned long long low,
                             int32_t *__restrict__ out) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz;
       e += int64_t(gridDim.x) * blockDim.x)
    out[e] = static_cast<int32_t>(keys[e] & low);
}

// Layout of census_relabel's scratch: degrees, ids, sorted degrees, the
// order and the ranks (n words each), the row lengths (n + 1 int64), the
// arc keys and their sorted copy (nnz u64 each), and the storage of the
// sorts and the scan.
struct RelabelTmp {
  size_t deg = 0, ids = 0, dsort = 0, perm = 0, rank = 0, lens = 0, keys = 0,
         keys2 = 0, work = 0, work_bytes = 0, total = 0;
  int lg = 1;  // bits of a vertex id
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

bool relabel_tmp(int32_t n, int64_t nnz, RelabelTmp *t) {
  const size_t nv = static_cast<size_t>(std::max<int32_t>(n, 1));
  const size_t na = static_cast<size_t>(std::max<int64_t>(nnz, 1));
  t->lg = 1;
  while (t->lg < 31 && (int64_t(1) << t->lg) < int64_t(n)) ++t->lg;
  size_t b1 = 0, b2 = 0, b3 = 0;
  if (hipcub::DeviceRadixSort::SortPairs(
          nullptr, b1, static_cast<const uint32_t *>(nullptr),
          static_cast<uint32_t *>(nullptr),
          static_cast<const uint32_t *>(nullptr),
          static_cast<uint32_t *>(nullptr), static_cast<int64_t>(nv), 0,
          32) != hipSuccess)
    return false;
  if (hipcub::DeviceRadixSort::SortKeys(
          nullptr, b2, static_cast<const unsigned long long *>(nullptr),
          static_cast<unsigned long long *>(nullptr),
          static_cast<int64_t>(na), 0, 2 * t->lg) != hipSuccess)
    return false;
  if (hipcub::DeviceScan::ExclusiveSum(
          nullptr, b3, static_cast<const int64_t *>(nullptr),
          static_cast<int64_t *>(nullptr),
          static_cast<int64_t>(nv + 1)) != hipSuccess)
    return false;
  size_t o = 0;
  t->deg = o;   o += align256(nv * 4);
  t->ids = o;   o += align256(nv * 4);
  t->dsort = o; o += align256(nv * 4);
  t->perm = o;  o += align256(nv * 4);
  t->rank = o;  o += align256(nv * 4);
  t->lens = o;  o += align256((nv + 1) * 8);
  t->keys = o;  o += align256(na * 8);
  t->keys2 = o; o += align256(na * 8);
  t->work = o;
  t->work_bytes = std::max<size_t>(std::max({b1, b2, b3}), 1);
  o += align256(t->work_bytes);
  t->total = o;
  return true;
}

// Workgroups of a class whose masks take `words` words per vertex: as many
// as the mask buffer holds slices for, at most c.groups.
int groups_for(const Census5Args &c, int words) {
  const int64_t g = c.mask_words / std::max(words, 1);
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g, c.groups)));
}

}  // namespace

size_t census5_group_bytes(int32_t n, int words) {
  return static_cast<size_t>(std::max<int32_t>(n, 1)) *
         (12 + 8 * static_cast<size_t>(words));
}

int census5_flat_launch(const Census5Args &c, int cus, hipStream_t s) {
  const int g = std::max(1, std::min(cus * 8, static_cast<int>(
                                                  (c.nnz + 255) / 256)));
  hipLaunchKernelGGL(c5_t_csr, dim3(g), dim3(256), 0, s, c);
  hipLaunchKernelGGL(c5_flat, dim3(g), dim3(256), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census5_table_count_launch(const Census5Args &c, hipStream_t s) {
  hipLaunchKernelGGL(c5_tab_count, dim3(std::max(1, c.groups)), dim3(kT), 0,
                     s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census5_classify_launch(const Census5Args &c, hipStream_t s) {
  const int gv = std::max(1, std::min(1024, (c.n + 255) / 256));
  hipLaunchKernelGGL(c5_class_count, dim3(gv), dim3(256), 0, s, c);
  hipLaunchKernelGGL(c5_class_scatter, dim3(gv), dim3(256), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census5_scan_bytes(int64_t count, size_t *bytes) {
  size_t b = 0;
  const hipError_t e = hipcub::DeviceScan::ExclusiveSum(
      nullptr, b, static_cast<const unsigned long long *>(nullptr),
      static_cast<unsigned long long *>(nullptr), count);
  *bytes = b;
  return e == hipSuccess ? 0 : -1;
}

int census5_scan(const unsigned long long *in, unsigned long long *out,
                 int64_t count, void *tmp, size_t tmp_bytes, hipStream_t s) {
  size_t b = tmp_bytes;
  return hipcub::DeviceScan::ExclusiveSum(tmp, b, in, out, count, s) ==
                 hipSuccess
             ? 0
             : -1;
}

size_t census_relabel_bytes(int32_t n, int64_t nnz) {
  RelabelTmp t;
  return relabel_tmp(n, nnz, &t) ? t.total : 0;
}

int census_relabel(const int64_t *rowptr, const int32_t *col, int32_t n,
                   int64_t nnz, int64_t *out_rowptr, int32_t *out_col,
                   void *tmp, size_t tmp_bytes, int cus, hipStream_t s) {
  if (n < 0) return -1;
  RelabelTmp t;
  if (!relabel_tmp(n, nnz, &t) || tmp_bytes < t.total) return -1;
  char *base = static_cast<char *>(tmp);
  auto *deg = reinterpret_cast<uint32_t *>(base + t.deg);
  auto *ids = reinterpret_cast<uint32_t *>(base + t.ids);
  auto *dsort = reinterpret_cast<uint32_t *>(base + t.dsort);
  auto *perm = reinterpret_cast<uint32_t *>(base + t.perm);
  auto *rank = reinterpret_cast<uint32_t *>(base + t.rank);
  auto *lens = reinterpret_cast<int64_t *>(base + t.lens);
  auto *keys = reinterpret_cast<unsigned long long *>(base + t.keys);
  auto *keys2 = reinterpret_cast<unsigned long long *>(base + t.keys2);
  void *work = base + t.work;
  const int gv = static_cast<int>(std::min<int64_t>(4096, n / 256 + 1));
  size_t b = t.work_bytes;
  if (n > 0) {
    hipLaunchKernelGGL(c5_degrees, dim3(gv), dim3(256), 0, s, rowptr, n, deg,
                       ids);
    // Radix sorts are stable: equal degrees keep id order (DAG order).
    if (hipcub::DeviceRadixSort::SortPairs(work, b, deg, dsort, ids, perm,
                                           static_cast<int64_t>(n), 0, 32,
                                           s) != hipSuccess)
      return -1;
  }
  hipLaunchKernelGGL(c5_ranks, dim3(gv), dim3(256), 0, s, perm, dsort, n,
                     rank, lens);
  b = t.work_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(work, b, lens, out_rowptr,
                                       static_cast<int64_t>(n) + 1,
                                       s) != hipSuccess)
    return -1;
  if (n > 0 && nnz > 0) {
    hipLaunchKernelGGL(c5_arc_keys, dim3(std::max(1, cus * 8)), dim3(256), 0,
                       s, rowptr, col, n, rank, t.lg, keys);
    b = t.work_bytes;
    if (hipcub::DeviceRadixSort::SortKeys(work, b, keys, keys2, nnz, 0,
                                          2 * t.lg, s) != hipSuccess)
      return -1;
    const int ga =
        static_cast<int>(std::min<int64_t>(8192, nnz / 256 + 1));
    hipLaunchKernelGGL(c5_arc_heads, dim3(ga), dim3(256), 0, s, keys2, nnz,
                       (1ull << t.lg) - 1ull, out_col);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census5_pairs_launch(const Census5Args &c, hipStream_t s) {
  hipLaunchKernelGGL(c5_anchor<1>, dim3(groups_for(c, 1)), dim3(kT), 0, s, c, 0);
  hipLaunchKernelGGL(c5_anchor<2>, dim3(groups_for(c, 2)), dim3(kT), 0, s, c, 1);
  hipLaunchKernelGGL(c5_anchor<4>, dim3(groups_for(c, 4)), dim3(kT), 0, s, c, 2);
  hipLaunchKernelGGL(c5_anchor<8>, dim3(groups_for(c, 8)), dim3(kT), 0, s, c, 3);
  hipLaunchKernelGGL(c5_anchor<0>, dim3(groups_for(c, c.gen_words)), dim3(kT),
                     0, s, c, 4);
  hipLaunchKernelGGL(c5_outnb<1>, dim3(std::max(1, c.groups)), dim3(kT), 0, s, c, 0);
  hipLaunchKernelGGL(c5_outnb<2>, dim3(std::max(1, c.groups)), dim3(kT), 0, s, c, 1);
  hipLaunchKernelGGL(c5_outnb<4>, dim3(std::max(1, c.groups)), dim3(kT), 0, s, c, 2);
  hipLaunchKernelGGL(c5_outnb<8>, dim3(std::max(1, c.groups)), dim3(kT), 0, s, c, 3);
  hipLaunchKernelGGL(c5_outnb<0>, dim3(std::max(1, c.groups)), dim3(kT), 0, s, c, 4);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

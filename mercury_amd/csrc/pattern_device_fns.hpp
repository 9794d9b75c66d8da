// P1..P8 device functions of the pattern kernels (pattern_kernels.hip,
// which describes the design).
#pragma once

#include <hip/hip_runtime.h>

#include "pattern_device.hpp"
#include "wave.hpp"

namespace mgpm {
namespace pdev {

struct PGraph {
  const int64_t *__restrict__ rowptr;
  const int32_t *__restrict__ col;
  __device__ __forceinline__ const int32_t *nb(int32_t v) const {
    return col + rowptr[v];
  }
  __device__ __forceinline__ int32_t deg(int32_t v) const {
    return static_cast<int32_t>(rowptr[v + 1] - rowptr[v]);
  }
};

// Exact membership of key in the sorted list a[0, n).
__device__ __forceinline__ bool bs_contains(const int32_t *__restrict__ a,
                                            int32_t n, int32_t key) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && a[lo] == key;
}

// First index i of the sorted list a[0, n) with a[i] > key (STRICT) or
// a[i] >= key.
template <bool STRICT>
__device__ __forceinline__ int32_t bound(const int32_t *__restrict__ a,
                                         int32_t n, int32_t key) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (STRICT ? a[mid] <= key : a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// floor(log2(n)) + 1 for n >= 1: the steps of a binary search over n.
__device__ __forceinline__ uint32_t search_steps(int32_t n) {
  return 32u - __clz(n | 1);
}

// Whether counting |A ∩ B| should walk B (each element searched in A)
// rather than A (each searched in B), A being a frontier list in the
// per-wave slab and B an adjacency list.  Each way costs ceil(n/64) wave
// steps of search_steps(other side) dependent loads; a search in the slab
// is charged 4 times one in the adjacency lists (the slab is far larger
// than the caches, a small graph's CSR is not), plus `extra` steps.
__device__ __forceinline__ bool walk_other_side(int32_t na, int32_t nb,
                                                uint32_t extra = 0) {
  const uint32_t walk_a = uint32_t((na + 63) >> 6) * search_steps(nb);
  const uint32_t walk_b = uint32_t((nb + 63) >> 6) * search_steps(na) + extra;
  return 4 * walk_b < walk_a;
}

// out = {x in A : x < ub, (x in B) == KEEP_IN}, order preserved.  Returns
// the count (wave-uniform).  `out` is wave-private global scratch.
template <bool KEEP_IN>
__device__ __forceinline__ int32_t wave_filter(const int32_t *A, int32_t nA,
                                               const int32_t *B, int32_t nB,
                                               int32_t ub, int32_t *out) {
  const int lane = lane64();
  int32_t count = 0;
  for (int32_t c = 0; c < nA; c += 64) {
    const int32_t i = c + lane;
    const bool valid = i < nA;
    const int32_t x = valid ? A[i] : 0;
    const bool below = valid && x < ub;
    const bool keep = below && (bs_contains(B, nB, x) == KEEP_IN);
    const unsigned long long m = __ballot(keep);
    if (keep) out[count + bits_below(m)] = x;
    count += __popcll(m);
    // A is sorted: once a valid element reaches ub, the rest do too.
    if (__ballot(valid && !below)) break;
  }
  return count;
}

// D(A,B) / D<(A,B,ub) of set_difference.cuh:129-237.
__device__ __forceinline__ int32_t wave_diff(const int32_t *A, int32_t nA,
                                             const int32_t *B, int32_t nB,
                                             int32_t ub, int32_t *out) {
  return wave_filter<false>(A, nA, B, nB, ub, out);
}

// I(A,B) / I<(A,B,ub) of set_intersect.cuh:73-193: walk the shorter list,
// search the longer one.
__device__ __forceinline__ int32_t wave_inter(const int32_t *A, int32_t nA,
                                              const int32_t *B, int32_t nB,
                                              int32_t ub, int32_t *out) {
  return nA <= nB ? wave_filter<true>(A, nA, B, nB, ub, out)
                  : wave_filter<true>(B, nB, A, nA, ub, out);
}

constexpr int32_t kNoUb = 0x7fffffff;

enum PairMode : int {
  kNonAdj = 0,     // x !~ y
  kNonAdjYGt = 1,  // y > x and x !~ y
  kAdjYGt = 2,     // y > x and x ~ y
  kAdj = 3,        // x ~ y
};

// Number of pairs (x in X, y in Y) satisfying MODE; per-lane partial.
// The outer loop runs over the shorter list (uniform element, its
// adjacency list searched by all lanes), the lanes over the longer one;
// only outer indices sub, sub + nsub, ... are taken (one share of a split
// task).
template <int MODE>
__device__ __forceinline__ uint64_t pair_count(const PGraph &g,
                                               const int32_t *X, int32_t nx,
                                               const int32_t *Y, int32_t ny,
                                               int32_t sub = 0,
                                               int32_t nsub = 1) {
  const int lane = lane64();
  uint64_t c = 0;
  if (nx == 0 || ny == 0) return 0;
  const bool outer_x = nx <= ny;
  const int32_t *O = outer_x ? X : Y;
  const int32_t *I = outer_x ? Y : X;
  const int32_t no = outer_x ? nx : ny, ni = outer_x ? ny : nx;
  for (int32_t a = sub; a < no; a += nsub) {
    const int32_t o = O[a];
    const int32_t *No = g.nb(o);
    const int32_t dO = g.deg(o);
    constexpr bool kOrdered = MODE == kNonAdjYGt || MODE == kAdjYGt;
    constexpr bool kCountAdj = MODE == kAdjYGt || MODE == kAdj;
    if (!walk_other_side(ni, dO, MODE == kNonAdjYGt ? search_steps(ni) : 0)) {
      for (int32_t b = lane; b < ni; b += 64) {
        const int32_t in = I[b];
        const int32_t x = outer_x ? o : in, y = outer_x ? in : o;
        if (kOrdered && !(y > x)) continue;
        const bool adj = bs_contains(No, dO, in);
        c += kCountAdj ? adj : !adj;
      }
      continue;
    }
    // Walk N(o) instead: adjacent partners are z in N(o) ∩ I (above o when
    // o is x, below o when o is y); non-adjacent ones are the eligible
    // inner elements minus those.
    uint32_t adj = 0;
    for (int32_t b = lane; b < dO; b += 64) {
      const int32_t z = No[b];
      const bool eligible = !kOrdered || (outer_x ? z > o : z < o);
      adj += eligible && bs_contains(I, ni, z);
    }
    if (kCountAdj) {
      c += adj;
    } else {
      int32_t total = ni;
      if (MODE != kNonAdj)
        total = outer_x ? ni - bound<true>(I, ni, o) : bound<false>(I, ni, o);
      // total - adj, split over the lanes; the wave sum is exact mod 2^64.
      if (lane == 0) c += static_cast<uint64_t>(total);
      c -= adj;
    }
  }
  return c;
}

// Per-wave scratch: six frontier lists and two per-index arrays.
struct Slab {
  int32_t *L[6];
  uint32_t *F;    // frequencies (lane-owned: index i by lane i mod 64)
  uint32_t *CNT;  // per-index counts written by lane 0
  __device__ Slab(int32_t *base, int64_t cap) {
    for (int k = 0; k < 6; ++k) L[k] = base + k * cap;
    F = reinterpret_cast<uint32_t *>(base + 6 * cap);
    CNT = reinterpret_cast<uint32_t *>(base + 7 * cap);
  }
};

__device__ __forceinline__ void zero_lane_owned(uint32_t *a, int32_t n) {
  for (int32_t i = lane64(); i < n; i += 64) a[i] = 0;
}

// Builds both sides of "L split by adjacency to v": NON = {x in L : x !~ v}
// and ADJ = {x in L : x ~ v} (either may be null), and adds x !~ v to the
// lane-owned F[] when F is given.  Returns the two counts.
__device__ __forceinline__ void split_by(const PGraph &g, int32_t v,
                                         const int32_t *L, int32_t n,
                                         int32_t *NON, int32_t &n_non,
                                         int32_t *ADJ, int32_t &n_adj,
                                         uint32_t *F) {
  const int lane = lane64();
  const int32_t *Nv = g.nb(v);
  const int32_t dv = g.deg(v);
  n_non = 0;
  n_adj = 0;
  for (int32_t c = 0; c < n; c += 64) {
    const int32_t i = c + lane;
    const bool valid = i < n;
    const int32_t x = valid ? L[i] : 0;
    const bool adj = valid && bs_contains(Nv, dv, x);
    const bool non = valid && !adj;
    if (F && non) F[i] += 1;
    const unsigned long long mn = __ballot(non), ma = __ballot(adj);
    if (NON && non) NON[n_non + bits_below(mn)] = x;
    if (ADJ && adj) ADJ[n_adj + bits_below(ma)] = x;
    n_non += __popcll(mn);
    n_adj += __popcll(ma);
  }
}

// Number of x in the sorted list L[0, upto) with (x ~ v) == ADJ, as a
// per-lane partial (the lanes' values may be negative; their sum is exact).
template <bool ADJ>
__device__ __forceinline__ int32_t count_prefix(const PGraph &g, int32_t v,
                                                const int32_t *L,
                                                int32_t upto) {
  const int32_t *Nv = g.nb(v);
  const int32_t dv = g.deg(v);
  const int lane = lane64();
  int32_t c = 0;
  if (!walk_other_side(upto, dv)) {
    for (int32_t i = lane; i < upto; i += 64)
      c += bs_contains(Nv, dv, L[i]) == ADJ;
    return c;
  }
  for (int32_t i = lane; i < dv; i += 64) c += bs_contains(L, upto, Nv[i]);
  if (ADJ) return c;
  return (lane == 0 ? upto : 0) - c;
}

// Number of x in L[0, n) not adjacent to v (wave-uniform total).
__device__ __forceinline__ uint32_t count_nonadj(const PGraph &g, int32_t v,
                                                 const int32_t *L, int32_t n) {
  return static_cast<uint32_t>(wave_sum_i32(count_prefix<false>(g, v, L, n)));
}

// |N(v) ∩ L[0, n)| for a sorted list L (wave-uniform), and with ABOVE also
// the part above v in *above.  The shorter of the two lists is walked and
// each element searched in the other.
template <bool ABOVE>
__device__ __forceinline__ uint32_t adj_in(const PGraph &g, int32_t v,
                                           const int32_t *L, int32_t n,
                                           uint32_t *above = nullptr) {
  const int32_t *Nv = g.nb(v);
  const int32_t dv = g.deg(v);
  const bool walk_n = dv <= n;
  const int32_t *W = walk_n ? Nv : L, *S = walk_n ? L : Nv;
  const int32_t nw = walk_n ? dv : n, ns = walk_n ? n : dv;
  int c = 0, u = 0;
  for (int32_t i = lane64(); i < nw; i += 64) {
    const int32_t x = W[i];
    const bool in = bs_contains(S, ns, x);
    c += in;
    if (ABOVE) u += in && x > v;
  }
  if (ABOVE) *above = static_cast<uint32_t>(wave_sum_i32(u));
  return static_cast<uint32_t>(wave_sum_i32(c));
}

// ------------------------------------------------------------ P1 vertex

// P1_frequency_count (pattern_kernels.cuh:68-133) for vertex v0:
// c1 += sum_k cnt[k] f[k] with cnt[k] = #{i < k : N0[i] !~ N0[k]} and
// f[k] = #{j > k : N0[k] !~ N0[j]}.  Unsplit (nsub == 1): one pass over
// the pairs i < j, f kept lane-owned.  Split: share `sub` takes the k with
// k = sub (mod nsub) and gets both terms of each from one pass over N0.
__device__ void p1_vertex(const PGraph &g, int32_t v0, int32_t sub,
                          int32_t nsub, Slab &s, uint64_t &c1) {
  const int lane = lane64();
  const int32_t d0 = g.deg(v0);
  const int32_t *N0 = g.nb(v0);
  if (nsub > 1) {
    for (int32_t k = sub; k < d0; k += nsub) {
      const int32_t y = N0[k];
      // N0[k] itself is not in N(y) (no self-loops), so it splits cleanly.
      const uint32_t before = count_nonadj(g, y, N0, k);
      const uint32_t after = count_nonadj(g, y, N0 + k + 1, d0 - k - 1);
      if (lane == 0) c1 += static_cast<uint64_t>(before) * after;
    }
    return;
  }
  zero_lane_owned(s.F, d0);
  for (int32_t j = 0; j < d0; ++j) {
    const int32_t y = N0[j];
    const int32_t *Ny = g.nb(y);
    const int32_t dy = g.deg(y);
    uint32_t cnt = 0;
    for (int32_t i = lane; i < j; i += 64) {
      const bool non = !bs_contains(Ny, dy, N0[i]);
      s.F[i] += non;
      cnt += non;
    }
    cnt = static_cast<uint32_t>(wave_sum_i32(static_cast<int>(cnt)));
    if (lane == 0) s.CNT[j] = cnt;
  }
  wave_sync();
  for (int32_t k = lane; k < d0; k += 64)
    c1 += static_cast<uint64_t>(s.CNT[k]) * s.F[k];
  wave_sync();
}

// -------------------------------------------------------------- edges

// One share of the task edge (v0, v1): outer loops take the indices
// sub, sub + nsub, ...; nsub == 1 is the whole edge.  Per-index frequency
// factors that the reference accumulates over the outer loop are computed
// per index here, so shares of one edge need no common state.
template <int P>
__device__ void edge_task(const PGraph &g, int32_t v0, int32_t v1,
                          int32_t sub, int32_t nsub, Slab &s, uint64_t &c1,
                          uint64_t &c2) {
  const int lane = lane64();
  const int32_t *N0 = g.nb(v0), *N1 = g.nb(v1);
  const int32_t d0 = g.deg(v0), d1 = g.deg(v1);
  int32_t **L = s.L;
  if constexpr (P == 1) {  // P1_count_correction :6-66
    if (v0 == v1) return;
    const int32_t na = wave_diff(N0, d0, N1, d1, v1, L[0]);
    const int32_t nb = wave_inter(N0, d0, N1, d1, v1, L[1]);
    wave_sync();
    c2 += pair_count<kNonAdjYGt>(g, L[1], nb, L[0], na, sub, nsub);
  } else if constexpr (P == 2 || P == 8) {  // :135-243, :461-572
    // The reference's S - C for one task edge is, over z in L1 = D(N1,N0)
    // with A_z = L0 ∩ N(z), the non-adjacent pairs (P2, L0 = D(N0,N1)) or
    // the adjacent pairs (P8, L0 = I(N0,N1)) of L0 \ A_z.  Expanding the
    // edges of L0 \ A_z:  E(L0 \ A_z) = E(L0) - sum_{a in A_z} d(a) + E(A_z)
    // with d(a) = |N(a) ∩ L0|, and sum_z sum_{a in A_z} d(a) =
    // sum_a d(a) g(a), g(a) = |N(a) ∩ L1|.  So per a in L0: n1 e+(a) and
    // d(a) g(a) (e+ = the part of N(a) ∩ L0 above a, counting E(L0) once);
    // per z in L1: C(|L0| - |A_z|, 2) (P2) and E(A_z).  Every term is an
    // integer per element, so shares need no common state.
    const int32_t n0 = P == 2 ? wave_diff(N0, d0, N1, d1, kNoUb, L[0])
                              : wave_inter(N0, d0, N1, d1, kNoUb, L[0]);
    const int32_t n1 = wave_diff(N1, d1, N0, d0, kNoUb, L[1]);
    wave_sync();
    for (int32_t i = sub; i < n0; i += nsub) {
      const int32_t a = L[0][i];
      uint32_t e = 0;
      const uint32_t d = adj_in<true>(g, a, L[0], n0, &e);
      const uint32_t gz = adj_in<false>(g, a, L[1], n1);
      if (lane == 0) {
        const uint64_t ne = static_cast<uint64_t>(n1) * e;
        const uint64_t dg = static_cast<uint64_t>(d) * gz;
        if (P == 2) { c2 += ne; c1 += dg; }
        else { c1 += ne; c2 += dg; }
      }
    }
    for (int32_t i = sub; i < n1; i += nsub) {
      const int32_t z = L[1][i];
      const int32_t k = wave_inter(g.nb(z), g.deg(z), L[0], n0, kNoUb, L[2]);
      wave_sync();
      const uint64_t ea = pair_count<kAdjYGt>(g, L[2], k, L[2], k);
      if (P == 2) {
        if (lane == 0) {
          const uint64_t r = static_cast<uint64_t>(n0 - k);
          c1 += r * (r - (r > 0 ? 1 : 0)) / 2;
        }
        c2 += ea;
      } else {
        c1 += ea;
      }
      wave_sync();
    }
  } else if constexpr (P == 3) {  // :245-350
    // The reference's S - C for one task edge is the number of independent
    // 3-subsets of L0 = D(N0,N1).  By
    // inclusion-exclusion over the three possible edges of a triple:
    //   indep3 = C(n,3) - (n-2) E + sum_v C(d_v, 2) - T
    // with d_v = |N(v) ∩ L0|, E the edges inside L0 and T its triangles.
    // E and T are counted once each by orientation (u > v), so every term
    // is an integer per vertex v and the shares of one edge need no common
    // state: c1 takes the added terms, c2 the subtracted ones.
    const int32_t n0 = wave_diff(N0, d0, N1, d1, kNoUb, L[0]);
    wave_sync();
    if (sub == 0 && lane == 0 && n0 >= 3) {
      const uint64_t n = static_cast<uint64_t>(n0);
      c1 += n * (n - 1) * (n - 2) / 6;
    }
    for (int32_t i = sub; i < n0; i += nsub) {
      const int32_t v = L[0][i];
      // M = N(v) ∩ L0 (sorted), and M+ = its part above v.
      const int32_t dm = wave_inter(g.nb(v), g.deg(v), L[0], n0, kNoUb, L[1]);
      wave_sync();
      const int32_t pos = bound<true>(L[1], dm, v);
      const int32_t up = dm - pos;
      if (lane == 0) {
        const uint64_t d = static_cast<uint64_t>(dm);
        c1 += d * (d - (d > 0 ? 1 : 0)) / 2;                       // C(d_v, 2)
        c2 += static_cast<uint64_t>(n0 - 2) * static_cast<uint64_t>(up);  // (n-2) e_v+
      }
      // Triangles v < a < b inside L0: adjacent pairs within M+.
      c2 += pair_count<kAdjYGt>(g, L[1] + pos, up, L[1] + pos, up);
      wave_sync();
    }
  } else if constexpr (P == 7) {  // :352-459
    // Independent 3-subsets of L0 = I(N0,N1) by the reference's triple
    // machinery (small dense lists: cheaper than the counting of P3).
    const int32_t n0 = wave_inter(N0, d0, N1, d1, kNoUb, L[0]);
    wave_sync();
    for (int32_t i = sub; i < n0; i += nsub) {
      const int32_t y = L[0][i];
      const int32_t *Ny = g.nb(y);
      const int32_t dy = g.deg(y);
      // A = D<(L0, N(y), y) (indices < i, not adjacent), B = I(L0, N(y));
      // c2 over pairs of B x A; c1 += |A| * #{j > i : L0[j] !~ y}.
      int32_t na = 0, nb = 0, nafter = 0;
      for (int32_t c = 0; c < n0; c += 64) {
        const int32_t j = c + lane;
        const bool valid = j < n0;
        const int32_t x = valid ? L[0][j] : 0;
        const bool adj = valid && bs_contains(Ny, dy, x);
        const bool inA = valid && j < i && !adj;
        const unsigned long long ma = __ballot(inA), mb = __ballot(adj);
        const unsigned long long mf = __ballot(valid && j > i && !adj);
        if (inA) L[1][na + bits_below(ma)] = x;
        if (adj) L[2][nb + bits_below(mb)] = x;
        na += __popcll(ma);
        nb += __popcll(mb);
        nafter += __popcll(mf);
      }
      if (lane == 0) c1 += static_cast<uint64_t>(na) * uint32_t(nafter);
      wave_sync();
      c2 += pair_count<kNonAdjYGt>(g, L[2], nb, L[1], na);
      wave_sync();
    }
  } else if constexpr (P == 5) {  // :574-734
    // S - C = sum over c in L0 = I(N0,N1) of the non-adjacent pairs of
    // (L1 \ A_c) x (L2 \ B_c), L1 = D(N0,N1), L2 = D(N1,N0), A_c = L1 ∩ N(c),
    // B_c = L2 ∩ N(c).  With E(X,Y) the cross edges,
    //   E(L1\A, L2\B) = E(L1,L2) - E(A,L2) - E(L1,B) + E(A,B),
    // and summed over c: sum_c E(A_c,L2) = sum_{a in L1} q(a) d2(a),
    // sum_c E(L1,B_c) = sum_{b in L2} q(b) d1(b) (q = |N(.) ∩ L0|,
    // d1/d2 = |N(.) ∩ L1/L2|), E(L1,L2) = sum_a d2(a).  Per a in L1:
    // (q(a) - n0) d2(a); per b in L2: q(b) d1(b); per c in L0:
    // |L1\A_c| |L2\B_c| - E(A_c,B_c).
    const int32_t n0 = wave_inter(N0, d0, N1, d1, kNoUb, L[0]);
    const int32_t n1 = wave_diff(N0, d0, N1, d1, kNoUb, L[1]);
    const int32_t n2 = wave_diff(N1, d1, N0, d0, kNoUb, L[2]);
    wave_sync();
    for (int32_t i = sub; i < n1; i += nsub) {
      const int32_t a = L[1][i];
      const uint32_t q = adj_in<false>(g, a, L[0], n0);
      const uint32_t d2 = adj_in<false>(g, a, L[2], n2);
      if (lane == 0) {
        c1 += static_cast<uint64_t>(q) * d2;
        c2 += static_cast<uint64_t>(n0) * d2;
      }
    }
    for (int32_t i = sub; i < n2; i += nsub) {
      const int32_t b = L[2][i];
      const uint32_t q = adj_in<false>(g, b, L[0], n0);
      const uint32_t d1v = adj_in<false>(g, b, L[1], n1);
      if (lane == 0) c1 += static_cast<uint64_t>(q) * d1v;
    }
    for (int32_t i = sub; i < n0; i += nsub) {
      const int32_t c = L[0][i];
      const int32_t ka = wave_inter(g.nb(c), g.deg(c), L[1], n1, kNoUb, L[3]);
      const int32_t kb = wave_inter(g.nb(c), g.deg(c), L[2], n2, kNoUb, L[4]);
      wave_sync();
      if (lane == 0)
        c1 += static_cast<uint64_t>(n1 - ka) * static_cast<uint64_t>(n2 - kb);
      c2 += pair_count<kAdj>(g, L[3], ka, L[4], kb);
      wave_sync();
    }
  } else if constexpr (P == 6) {  // :736-853
    // sum over c in L0 = I(N0,N1) of the non-adjacent pairs of
    // (L0 \ A_c) x (L1 \ B_c), L1 = D(N1,N0), A_c = L0 ∩ N(c), B_c = L1 ∩ N(c)
    // (c itself stays in L0 \ A_c, as D(L0, N(c)) keeps it).  Expanding the
    // cross edges as for P5: per c in L0: |L0\A_c| |L1\B_c| + |A_c| |B_c|
    // - n0 |B_c| - E(A_c,B_c); per b in L1: |N(b) ∩ L0|^2.  P6 reports c1
    // alone (sm_multigpu.cu:209-232), so the subtracted terms come off c1
    // (exact modulo 2^64).
    const int32_t n0 = wave_inter(N0, d0, N1, d1, kNoUb, L[0]);
    const int32_t n1 = wave_diff(N1, d1, N0, d0, kNoUb, L[1]);
    wave_sync();
    for (int32_t i = sub; i < n0; i += nsub) {
      const int32_t c = L[0][i];
      const int32_t ka = wave_inter(g.nb(c), g.deg(c), L[0], n0, kNoUb, L[2]);
      const int32_t kb = wave_inter(g.nb(c), g.deg(c), L[1], n1, kNoUb, L[3]);
      wave_sync();
      if (lane == 0) {
        c1 += static_cast<uint64_t>(n0 - ka) * static_cast<uint64_t>(n1 - kb) +
              static_cast<uint64_t>(ka) * static_cast<uint64_t>(kb);
        c1 -= static_cast<uint64_t>(n0) * static_cast<uint64_t>(kb);
      }
      c1 -= pair_count<kAdj>(g, L[2], ka, L[3], kb);
      wave_sync();
    }
    for (int32_t i = sub; i < n1; i += nsub) {
      const uint64_t q = adj_in<false>(g, L[1][i], L[0], n0);
      if (lane == 0) c1 += q * q;
    }
  } else if constexpr (P == 4) {  // :855-963, v1 < v0 only
    if (v1 >= v0) return;
    const int32_t n0 = wave_diff(N0, d0, N1, d1, kNoUb, L[0]);
    const int32_t n1 = wave_diff(N1, d1, N0, d0, kNoUb, L[1]);
    wave_sync();
    c1 += pair_count<kNonAdj>(g, L[0], n0, L[1], n1, sub, nsub);
  }
  wave_sync();
}

__device__ __forceinline__ void flush2(uint64_t c1, uint64_t c2,
                                       unsigned long long *counters) {
  c1 = wave_sum_u64(c1);
  c2 = wave_sum_u64(c2);
  if (lane64() == 0) {
    if (c1) atomicAdd(&counters[0], static_cast<unsigned long long>(c1));
    if (c2) atomicAdd(&counters[1], static_cast<unsigned long long>(c2));
  }
}

}  // namespace pdev
}  // namespace mgpm

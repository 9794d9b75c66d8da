// Subgraph census by triangle enumeration over the degree-oriented DAG,
// and a 4-cycle count: the terms from which P1 (induced claws) and P4
// (induced paths on 4 vertices) follow in closed form.
//
// One wave per DAG arc x->y (arcs handed out by an atomic cursor, so the
// arcs of a hub row spread over many waves):
//   * C = N+(x) ∩ N+(y): the lanes take the shorter list and binary-search
//     the longer; every hit z is a triangle {x, y, z}, seen exactly once
//     (x < y < z in DAG order).  Its three edges gain one triangle each in
//     t[] (per DAG arc: x->y, x->z, y->z), and the triangle adds
//     d(x) + d(y) + d(z) - 6 to the degree term;
//   * K4s through the triangle set of the arc: ordered pairs (z, q) of C
//     with q ∈ N+(z) -- each K4 {x, y, z, q} once -- by pairs for small C,
//     by per-z intersections for large C.
// A second pass over the DAG arcs sums C(t_e, 2), C(t_e, 3) and C(d_v, 3).
// A third pass (census_k5, P7 and P8 only) takes the 4-cliques on every
// triangle and the 5-cliques.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "census.hpp"
#include "choose.hpp"
#include "wave.hpp"

namespace mgpm {

namespace {

constexpr int kWaves = 4;
constexpr int kPairMax = 64;  // |C| up to this: K4 test by flattened pairs

// First index in sorted list[0, len) with list[i] >= key.
__device__ __forceinline__ int lower_bound_g(const int32_t *__restrict__ list,
                                             int len, int32_t key) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (list[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool contains_g(const int32_t *__restrict__ list,
                                           int len, int32_t key) {
  const int i = lower_bound_g(list, len, key);
  return i < len && list[i] == key;
}

// How many of K keys (flags ok) lie in the sorted list[0, len), len >= 1:
// K branch-free searches for the last element <= key run in lockstep (the
// window shrinks by the same half for every key), so their loads overlap.
// For K keys, the position of the last element <= key in the sorted
// list[0, len), len >= 1 (0 if every element is larger): the searches of
// found_in, returning positions.
template <int K>
__device__ __forceinline__ void last_le(const int32_t *__restrict__ list,
                                        int len, const int32_t (&key)[K],
                                        int (&b)[K]) {
#pragma unroll
  for (int j = 0; j < K; ++j) b[j] = 0;
  for (int n = len; n > 1;) {
    const int half = n >> 1;
    int32_t v[K];
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = list[b[j] + half];
#pragma unroll
    for (int j = 0; j < K; ++j) b[j] = v[j] <= key[j] ? b[j] + half : b[j];
    n -= half;
  }
}

template <int K>
__device__ __forceinline__ void found_in(const int32_t *__restrict__ list,
                                         int len, const int32_t (&key)[K],
                                         const bool (&ok)[K], bool (&found)[K]) {
  int b[K];
#pragma unroll
  for (int j = 0; j < K; ++j) b[j] = 0;
  for (int n = len; n > 1;) {
    const int half = n >> 1;
    int32_t v[K];
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] = list[b[j] + half];
#pragma unroll
    for (int j = 0; j < K; ++j) b[j] = v[j] <= key[j] ? b[j] + half : b[j];
    n -= half;
  }
#pragma unroll
  for (int j = 0; j < K; ++j) found[j] = ok[j] && list[b[j]] == key[j];
}

// found_in with a list per key (lengths >= 1): each key's window halves
// until it is one element; the loop runs as long as the longest list
// needs, keys whose window is done re-read their element.
template <int K>
__device__ __forceinline__ void found_in_each(const int32_t *const (&list)[K],
                                              const int (&len)[K],
                                              const int32_t (&key)[K],
                                              const bool (&ok)[K],
                                              bool (&found)[K]) {
  int b[K], n[K], nmax = 1;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    b[j] = 0;
    n[j] = len[j];
    nmax = n[j] > nmax ? n[j] : nmax;
  }
  for (; nmax > 1; nmax -= nmax >> 1) {
    int32_t v[K];
    int half[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      half[j] = n[j] >> 1;
      v[j] = list[j][b[j] + half[j]];
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      b[j] = v[j] <= key[j] ? b[j] + half[j] : b[j];
      n[j] -= half[j];
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) found[j] = ok[j] && list[j][b[j]] == key[j];
}

template <int K>
__device__ __forceinline__ int count_in(const int32_t *__restrict__ list,
                                        int len, const int32_t (&key)[K],
                                        const bool (&ok)[K]) {
  bool found[K];
  found_in<K>(list, len, key, ok, found);
  int hits = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) hits += found[j] ? 1 : 0;
  return hits;
}

__global__ void dag_sources(const int64_t *__restrict__ drp, int32_t n,
                            int32_t *__restrict__ dsrc) {
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n;
       v += int64_t(gridDim.x) * blockDim.x)
    for (int64_t a = drp[v]; a < drp[v + 1]; ++a)
      dsrc[a] = static_cast<int32_t>(v);
}

__global__ __launch_bounds__(kWaves * 64) void census_triangles(CensusArgs c) {
  const int lane = lane64();
  const int64_t wid = int64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6);
  int32_t *slab = c.slab + wid * c.slab_len;
  unsigned long long tri = 0, tdeg = 0, k4 = 0, crk = 0, bdeg = 0;
  for (;;) {
    uint32_t a = 0;
    if (lane == 0) a = atomicAdd(c.cursor, 1u);
    a = wave_first(a);
    if (static_cast<int64_t>(a) >= c.m) break;
    const int32_t x = c.dsrc[a], y = c.dcol[a];
    const int64_t bx = c.drp[x], by = c.drp[y];
    const int lx = static_cast<int>(c.drp[x + 1] - bx);
    const int ly = static_cast<int>(c.drp[y + 1] - by);
    // Stream the shorter list S, search the longer L.
    const bool xs = lx <= ly;
    const int64_t sb = xs ? bx : by, lb = xs ? by : bx;
    const int ls = xs ? lx : ly, ll = xs ? ly : lx;
    const int32_t *S = c.dcol + sb;
    const int32_t *L = c.dcol + lb;
    int cnt = 0;
    // 4 items of S per lane at a time, searched in L in lockstep; hits are
    // compacted in item order, so C stays sorted.
    for (int p0 = 0; ls > 0 && ll > 0 && p0 < ls; p0 += 4 * 64) {
      int32_t zk[4];
      int jk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = p0 + 64 * k + lane;
        zk[k] = S[p < ls ? p : ls - 1];
      }
      last_le<4>(L, ll, zk, jk);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = p0 + 64 * k + lane;
        const int32_t z = zk[k];
        const int j = jk[k];
        const bool hit = p < ls && L[j] == z;
        if (hit) {
          atomicAdd(&c.t[sb + p], 1u);
          atomicAdd(&c.t[lb + j], 1u);
          const unsigned long long dz = static_cast<unsigned long long>(c.deg[z]);
          tdeg += dz;
          if (c.full5) {
            // Degrees in a triangle are >= 2: every factor is >= 0.
            const unsigned long long ex = c.deg[x] - 2, ey = c.deg[y] - 2,
                                     ez = dz - 2;
            crk += choose2(ex) + choose2(ey) + choose2(ez);
            bdeg += ex * ey + ex * ez + ey * ez;
          }
        }
        const unsigned long long mask = __ballot(hit);
        if (hit) slab[cnt + bits_below(mask)] = z;
        cnt += __popcll(mask);
      }
    }
    if (cnt == 0) continue;
    if (lane == 0) {
      atomicAdd(&c.t[a], static_cast<unsigned>(cnt));
      tri += static_cast<unsigned long long>(cnt);
      tdeg += static_cast<unsigned long long>(cnt) *
              static_cast<unsigned long long>(c.deg[x] + c.deg[y] - 6);
    }
    if (cnt < 2) continue;
    wave_sync();  // the slab entries written by other lanes
    if (cnt <= kPairMax) {
      // Ordered pairs (i, j) of C, 4 per lane at a time: is C[j] in N+(C[i])?
      const int np = cnt * cnt;
      for (int p0 = 0; p0 < np; p0 += 4 * 64) {
        const int32_t *lst[4];
        int len[4];
        int32_t key[4];
        bool ok[4], hit[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int p = p0 + 64 * k + lane;
          const int pp = p < np ? p : 0;
          const int i = pp / cnt, j = pp - i * cnt;
          const int32_t z = slab[i];
          const int64_t bz = c.drp[z];
          len[k] = static_cast<int>(c.drp[z + 1] - bz);
          ok[k] = p < np && i != j && len[k] > 0;
          lst[k] = c.dcol + bz;
          if (len[k] < 1) {  // a one-element window at a valid address
            lst[k] = slab;
            len[k] = 1;
          }
          key[k] = slab[j];
        }
        found_in_each<4>(lst, len, key, ok, hit);
#pragma unroll
        for (int k = 0; k < 4; ++k) k4 += hit[k] ? 1 : 0;
      }
    } else {
      for (int i = 0; i < cnt; ++i) {
        const int32_t z = slab[i];
        const int64_t bz = c.drp[z];
        const int lz = static_cast<int>(c.drp[z + 1] - bz);
        const int32_t *Z = c.dcol + bz;
        // Intersect C with N+(z): stream the shorter, search the longer.
        if (cnt <= lz) {
          for (int p = lane; p < cnt; p += 64)
            if (contains_g(Z, lz, slab[p])) ++k4;
        } else {
          for (int p = lane; p < lz; p += 64)
            if (contains_g(slab, cnt, Z[p])) ++k4;
        }
      }
    }
    wave_sync();  // before the next arc overwrites the slab
  }
  tri = wave_sum_u64(tri);
  tdeg = wave_sum_u64(tdeg);
  k4 = wave_sum_u64(k4);
  crk = wave_sum_u64(crk);
  bdeg = wave_sum_u64(bdeg);
  if (lane == 0) {
    if (tri) atomicAdd(&c.acc[kCensusTriangles], tri);
    if (tdeg) atomicAdd(&c.acc[kCensusTriDegrees], tdeg);
    if (k4) atomicAdd(&c.acc[kCensusK4], k4);
    if (crk) atomicAdd(&c.acc[kCensusCricket], crk);
    if (bdeg) atomicAdd(&c.acc[kCensusBullDeg], bdeg);
  }
}

// Sums over vertices (C(d,3)) and over DAG arcs (C(t,2) and
// (d_x - t - 1)(d_y - t - 1)).
// With full5 also C(d,4), the fork and diamond-pendant sums per edge, and
// twice the triangles per vertex into tv.
__global__ void census_sums(CensusArgs c) {
  unsigned long long s3 = 0, s2 = 0, sp = 0, st = 0, s4 = 0, fk = 0, dp3 = 0;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < c.n;
       v += int64_t(gridDim.x) * blockDim.x) {
    const unsigned long long d = static_cast<unsigned long long>(c.deg[v]);
    s3 += choose3(d);
    if (c.full5) s4 += choose4(d);
    unsigned long long tsum = 0;
    for (int64_t a = c.drp[v]; a < c.drp[v + 1]; ++a) {
      const unsigned long long t = c.t[a];
      const int32_t y = c.dcol[a];
      const unsigned long long dy = static_cast<unsigned long long>(c.deg[y]);
      s2 += choose2(t);
      st += choose3(t);
      sp += (d - t - 1) * (dy - t - 1);
      if (c.full5) {
        // d, dy >= 1; t > 0 only when both are >= 2, so t (d - 2) is 0
        // whenever d - 2 would wrap.
        fk += choose2(d - 1) * (dy - 1) - t * (d - 2) +
              choose2(dy - 1) * (d - 1) - t * (dy - 2);
        dp3 += choose2(t) * (d + dy - 6);
        tsum += t;
        if (t) atomicAdd(&c.tv[y], t);
      }
    }
    if (c.full5 && tsum) atomicAdd(&c.tv[v], tsum);
  }
  s3 = wave_sum_u64(s3);
  s2 = wave_sum_u64(s2);
  sp = wave_sum_u64(sp);
  st = wave_sum_u64(st);
  s4 = wave_sum_u64(s4);
  fk = wave_sum_u64(fk);
  dp3 = wave_sum_u64(dp3);
  if (lane64() == 0) {
    if (s3) atomicAdd(&c.acc[kCensusDeg3], s3);
    if (s2) atomicAdd(&c.acc[kCensusEdgePairs], s2);
    if (sp) atomicAdd(&c.acc[kCensusPathPairs], sp);
    if (st) atomicAdd(&c.acc[kCensusEdgeTriples], st);
    if (s4) atomicAdd(&c.acc[kCensusDeg4], s4);
    if (fk) atomicAdd(&c.acc[kCensusFork], fk);
    if (dp3) atomicAdd(&c.acc[kCensusDp3], dp3);
  }
}

// Per-triangle 4-clique terms and 5-cliques, one wave per DAG arc x->y
// (arcs handed out by an atomic cursor).  C = N+(x) ∩ N+(y) is the set of
// triangles {x, y, z} seen at this arc (each triangle once), kept with the
// positions of x->z and y->z in the DAG column array.
//   * k_T = |N(x) ∩ N(y) ∩ N(z)|, the 4-cliques on the triangle, from the
//     full common neighbourhood F = N(x) ∩ N(y) (|F| must equal t_e; for
//     small F the membership tests of 64 triangles at a time are spread
//     over the lanes together); the triangle adds C(k_T, 2),
//     k_T (t_xy + t_xz + t_yz - 6) and k_T (d_x + d_y + d_z - 9);
//   * the DAG adjacency inside C as bit rows (bit j of row i: C[j] in
//     N+(C[i])): row i's bits are the 4-cliques {x, y, C[i], C[j]}, and
//     |row i ∩ row j| over the bits j of row i are the 5-cliques, each
//     once (x, y, C[i], C[j], C[k] in DAG order).
// Slab per wave: C, the two position arrays and the row-bit lists (max
// out-degree each), F (max degree), then the rows (max_out x
// ceil(max_out / 64) words).
constexpr int kSerialKt = 128;  // |F| up to this: one triangle per lane

__global__ __launch_bounds__(kWaves * 64) void census_k5(CensusArgs c) {
  __shared__ int s_off[kWaves][64];
  __shared__ int s_dz[kWaves][64];
  __shared__ int64_t s_base[kWaves][64];
  __shared__ unsigned s_hits[kWaves][64];
  const int lane = lane64();
  const int wv = threadIdx.x >> 6;
  const int64_t wid = int64_t(blockIdx.x) * kWaves + wv;
  const int64_t L = c.slab_len;
  int32_t *Cz = c.k5_slab + wid * c.k5_len;
  uint32_t *Cxz = reinterpret_cast<uint32_t *>(Cz + L);
  uint32_t *Cyz = reinterpret_cast<uint32_t *>(Cz + 2 * L);
  int32_t *ML = Cz + 3 * L;
  int32_t *F = Cz + 4 * L;
  unsigned long long *rows = reinterpret_cast<unsigned long long *>(
      Cz + ((4 * L + c.max_deg + 1) & ~int64_t(1)));
  unsigned long long kk = 0, tek = 0, tdk = 0, k4 = 0, k5 = 0, bad = 0,
                     dp2 = 0, gem = 0;
  for (;;) {
    uint32_t a = 0;
    if (lane == 0) a = atomicAdd(c.cursor + 2, 1u);
    a = wave_first(a);
    if (static_cast<int64_t>(a) >= c.m) break;
    const uint32_t ta = c.t[a];
    if (ta == 0) continue;  // no triangle on the edge
    const int32_t x = c.dsrc[a], y = c.dcol[a];
    const int64_t bx = c.drp[x], by = c.drp[y];
    const int lx = static_cast<int>(c.drp[x + 1] - bx);
    const int ly = static_cast<int>(c.drp[y + 1] - by);
    const bool xs = lx <= ly;
    const int64_t sb = xs ? bx : by, lb = xs ? by : bx;
    const int ls = xs ? lx : ly, ll = xs ? ly : lx;
    const int32_t *S = c.dcol + sb;
    const int32_t *Lg = c.dcol + lb;
    int cnt = 0;
    for (int p0 = 0; ls > 0 && ll > 0 && p0 < ls; p0 += 4 * 64) {
      int32_t zk[4];
      int jk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = p0 + 64 * k + lane;
        zk[k] = S[p < ls ? p : ls - 1];
      }
      last_le<4>(Lg, ll, zk, jk);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = p0 + 64 * k + lane;
        const int32_t z = zk[k];
        const int j = jk[k];
        const bool hit = p < ls && Lg[j] == z;
        const unsigned long long mask = __ballot(hit);
        if (hit) {
          const int o = cnt + bits_below(mask);
          const uint32_t is = static_cast<uint32_t>(sb + p);
          const uint32_t il = static_cast<uint32_t>(lb + j);
          Cz[o] = z;
          Cxz[o] = xs ? is : il;
          Cyz[o] = xs ? il : is;
        }
        cnt += __popcll(mask);
      }
    }
    if (cnt == 0) continue;  // x and y are not the lowest of any triangle
    // F = N(x) ∩ N(y) over the undirected lists.
    const int dx = c.deg[x], dy = c.deg[y];
    const bool wx = dx <= dy;
    const int32_t *W = c.col + c.rowptr[wx ? x : y];
    const int32_t *Sr = c.col + c.rowptr[wx ? y : x];
    const int nw = wx ? dx : dy, ns = wx ? dy : dx;
    int te = 0;
    for (int p0 = 0; nw > 0 && p0 < nw; p0 += 4 * 64) {  // 4 per lane, in order
      int32_t wk[4];
      bool ok[4], hk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = p0 + 64 * k + lane;
        ok[k] = p < nw;
        wk[k] = W[ok[k] ? p : nw - 1];
      }
      found_in<4>(Sr, ns, wk, ok, hk);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned long long mask = __ballot(hk[k]);
        if (hk[k]) F[te + bits_below(mask)] = wk[k];
        te += __popcll(mask);
      }
    }
    if (lane == 0 && static_cast<uint32_t>(te) != ta) ++bad;
    wave_sync();  // C and F written by other lanes
    if (c.full5) {
      // Per triangle {x, y, z} (each seen once, at its lowest arc): the
      // diamond-pendant and gem terms from its three edge counts (>= 1)
      // and degrees (>= 2).
      for (int i = lane; i < cnt; i += 64) {
        const unsigned long long txy = ta - 1, txz = c.t[Cxz[i]] - 1,
                                 tyz = c.t[Cyz[i]] - 1;
        const unsigned long long ez = c.deg[Cz[i]] - 2;
        dp2 += txy * ez + txz * static_cast<unsigned long long>(dy - 2) +
               tyz * static_cast<unsigned long long>(dx - 2);
        gem += txy * txz + txy * tyz + txz * tyz;
      }
    }
    const unsigned long long dxy = static_cast<unsigned long long>(dx + dy);
    if (te <= kSerialKt) {
      // 64 triangles at a time; their membership tests (the shorter of F
      // and N(z) searched in the other) flattened over the lanes, the hits
      // counted per triangle in LDS.
      for (int i0 = 0; i0 < cnt; i0 += 64) {
        const int i = i0 + lane;
        int len = 0, dz = 0;
        int64_t bz = 0;
        if (i < cnt) {
          const int32_t z = Cz[i];
          dz = c.deg[z];
          bz = c.rowptr[z];
          len = te <= dz ? te : dz;
        }
        const int incl = wave_incl_scan(len);
        const int total = __shfl(incl, 63, 64);
        s_off[wv][lane] = incl - len;
        s_dz[wv][lane] = dz;
        s_base[wv][lane] = bz;
        s_hits[wv][lane] = 0;
        wave_sync();  // the per-triangle LDS entries
        for (int p0 = 0; p0 < total; p0 += 2 * 64) {  // 2 tests per lane
          const int32_t *lst[2];
          int len[2], rk[2];
          int32_t key[2];
          bool ok[2], hit[2];
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            const int p = p0 + 64 * k + lane;
            ok[k] = p < total;
            const int pp = ok[k] ? p : 0;
            // r = the last triangle whose run starts at or before pp.
            int lo = 0, hi = 64;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (s_off[wv][mid] <= pp) lo = mid + 1;
              else hi = mid;
            }
            const int r = lo - 1;
            const int q = pp - s_off[wv][r];
            const int dr = s_dz[wv][r];
            const int32_t *Nz = c.col + s_base[wv][r];
            const bool fz = te <= dr;  // search F[q] in N(z), else Nz[q] in F
            lst[k] = fz ? Nz : F;
            len[k] = fz ? dr : te;
            key[k] = fz ? F[q] : Nz[q];
            rk[k] = r;
          }
          found_in_each<2>(lst, len, key, ok, hit);
#pragma unroll
          for (int k = 0; k < 2; ++k)
            if (hit[k]) atomicAdd(&s_hits[wv][rk[k]], 1u);
        }
        wave_sync();  // the hit counts
        const unsigned k = s_hits[wv][lane];
        if (k) {
          const unsigned long long kt = k;
          kk += kt * (kt - 1) / 2;
          tek += kt * (static_cast<unsigned long long>(ta) + c.t[Cxz[i]] +
                       c.t[Cyz[i]] - 6);
          tdk += kt * (dxy + static_cast<unsigned long long>(dz) - 9);
        }
        wave_sync();  // before the next 64 triangles overwrite the entries
      }
    } else {
      for (int i = 0; i < cnt; ++i) {
        const int32_t z = Cz[i];
        const int dz = c.deg[z];
        const int32_t *Nz = c.col + c.rowptr[z];
        // The shorter of F and N(z) searched in the other, 4 keys per lane
        // at a time (idle keys repeat a valid one and are not counted).
        const bool fz = te <= dz;
        const int32_t *keys = fz ? F : Nz;
        const int32_t *list = fz ? Nz : F;
        const int nk = fz ? te : dz, nl = fz ? dz : te;
        int part = 0;
        for (int q0 = 0; q0 < nk; q0 += 4 * 64) {
          int32_t key[4];
          bool ok[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int q = q0 + 64 * j + lane;
            ok[j] = q < nk;
            key[j] = keys[ok[j] ? q : nk - 1];
          }
          part += count_in<4>(list, nl, key, ok);
        }
        const unsigned k = static_cast<unsigned>(wave_sum_i32(part));
        if (lane == 0 && k) {
          const unsigned long long kt = k;
          kk += kt * (kt - 1) / 2;
          tek += kt * (static_cast<unsigned long long>(ta) + c.t[Cxz[i]] +
                       c.t[Cyz[i]] - 6);
          tdk += kt * (dxy + static_cast<unsigned long long>(dz) - 9);
        }
      }
    }
    if (cnt >= 2) {
      const int R = (cnt + 63) >> 6;
      if (cnt <= 32) {
        // 64 / cnt rows per ballot: lane l tests C[l % cnt] against row
        // i0 + l / cnt, and row i0 + r is bits [r cnt, (r + 1) cnt).
        const int rpb = 64 / cnt;
        const int li = lane / cnt, j = lane - li * cnt;
        for (int i0 = 0; i0 < cnt; i0 += rpb) {
          const int i = i0 + li;
          bool adj = false;
          if (li < rpb && i < cnt) {
            const int32_t zi = Cz[i];
            const int64_t bi = c.drp[zi];
            const int di = static_cast<int>(c.drp[zi + 1] - bi);
            adj = contains_g(c.dcol + bi, di, Cz[j]);
          }
          const unsigned long long mask = __ballot(adj);
          if (lane < rpb && i0 + lane < cnt)
            rows[i0 + lane] = (mask >> (lane * cnt)) & ((1ull << cnt) - 1);
        }
      } else {
        for (int i = 0; i < cnt; ++i) {
          const int32_t zi = Cz[i];
          const int64_t bi = c.drp[zi];
          const int di = static_cast<int>(c.drp[zi + 1] - bi);
          const int32_t *Ni = c.dcol + bi;
          for (int w0 = 0; w0 < R; w0 += 4) {  // 4 row words per search
            int32_t key[4];
            bool ok[4], adj[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int j = (w0 + k) * 64 + lane;
              ok[k] = w0 + k < R && j < cnt;
              key[k] = Cz[ok[k] ? j : 0];
              adj[k] = false;
            }
            if (di > 0) found_in<4>(Ni, di, key, ok, adj);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              if (w0 + k >= R) break;  // uniform
              const unsigned long long mask = __ballot(adj[k]);
              if (lane == 0) rows[int64_t(i) * R + w0 + k] = mask;
            }
          }
        }
      }
      wave_sync();  // rows written by other lanes
      if (R == 1) {
        for (int i = lane; i < cnt; i += 64) {
          const unsigned long long ri = rows[i];
          k4 += static_cast<unsigned long long>(__popcll(ri));
          for (unsigned long long b = ri; b; b &= b - 1)
            k5 += static_cast<unsigned long long>(
                __popcll(ri & rows[__builtin_ctzll(b)]));
        }
      } else {
        for (int i = 0; i < cnt; ++i) {
          const unsigned long long *ri = rows + int64_t(i) * R;
          int nj = 0;  // the bits of row i, listed in ML
          for (int w0 = 0; w0 < R; w0 += 64) {
            const int w = w0 + lane;
            const unsigned long long word = w < R ? ri[w] : 0ull;
            const int pc = __popcll(word);
            const int incl = wave_incl_scan(pc);
            int o = nj + incl - pc;
            for (unsigned long long b = word; b; b &= b - 1)
              ML[o++] = w * 64 + __builtin_ctzll(b);
            nj += __shfl(incl, 63, 64);
          }
          wave_sync();  // ML written by other lanes
          if (lane == 0) k4 += static_cast<unsigned long long>(nj);
          for (int p = lane; p < nj; p += 64) {
            const unsigned long long *rj = rows + int64_t(ML[p]) * R;
            unsigned long long s5 = 0;
            for (int w = 0; w < R; ++w)
              s5 += static_cast<unsigned long long>(__popcll(ri[w] & rj[w]));
            k5 += s5;
          }
          wave_sync();  // before ML is overwritten
        }
      }
    }
    wave_sync();  // before the next arc overwrites the slab
  }
  kk = wave_sum_u64(kk);
  tek = wave_sum_u64(tek);
  tdk = wave_sum_u64(tdk);
  k4 = wave_sum_u64(k4);
  k5 = wave_sum_u64(k5);
  bad = wave_sum_u64(bad);
  dp2 = wave_sum_u64(dp2);
  gem = wave_sum_u64(gem);
  if (lane == 0) {
    if (dp2) atomicAdd(&c.acc[kCensusDp2], dp2);
    if (gem) atomicAdd(&c.acc[kCensusGem], gem);
    if (kk) atomicAdd(&c.acc[kCensusTriPairs], kk);
    if (tek) atomicAdd(&c.acc[kCensusTriEdgeK], tek);
    if (tdk) atomicAdd(&c.acc[kCensusTriDegK], tdk);
    if (k4) atomicAdd(&c.acc[kCensusK4Rows], k4);
    if (k5) atomicAdd(&c.acc[kCensusK5], k5);
    if (bad) atomicAdd(&c.acc[kCensusMismatch], bad);
  }
}

// 4-cycles as subgraphs, each counted once at its highest-ranked vertex v
// (rank: degree, then id -- the DAG order): for every neighbour u of v
// ranked below v and every neighbour w of u ranked below v, count the
// wedge v-u-w at w; a cycle v-u-w-u' is one pair of wedges at w, so v adds
// sum_w C(wedges(w), 2).  One workgroup per v (vertices handed out by an
// atomic cursor), with its own dense counter array over all vertices
// (left zeroed) and the list of the w it touched.
constexpr int kC4Threads = 256;

__global__ __launch_bounds__(kC4Threads) void census_c4(CensusArgs c) {
  uint32_t *cnt = c.wedges + size_t(blockIdx.x) * size_t(c.n);
  int32_t *touched = c.touched + size_t(blockIdx.x) * size_t(c.n);
  __shared__ unsigned s_v, s_nt;
  __shared__ unsigned long long s_acc[kC4Threads / 64],
      s_ban[kC4Threads / 64], s_tq[kC4Threads / 64];
  const int lane = lane64(), wv = threadIdx.x >> 6;
  constexpr int nwv = kC4Threads / 64;
  unsigned long long acc = 0, ban = 0, tq = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      s_v = atomicAdd(c.cursor + 1, 1u);
      s_nt = 0;
    }
    __syncthreads();
    const unsigned v = s_v;
    if (v >= static_cast<unsigned>(c.n)) break;
    // u below v: the first d(v) - d+(v) entries of v's sorted row; w below
    // v: the entries of u's row before v (the graph is relabelled in DAG
    // order, so rank is id).
    const int64_t ev = c.rowptr[v + 1] - (c.drp[v + 1] - c.drp[v]);
    for (int64_t i = c.rowptr[v] + wv; i < ev; i += nwv) {
      const int32_t u = c.col[i];
      const int64_t bu = c.rowptr[u];
      const int64_t eu = wave_first_greater(
          c.col, static_cast<int32_t>(v) - 1, bu, c.rowptr[u + 1]);
      for (int64_t j = bu + lane; j < eu; j += 64) {
        const int32_t w = c.col[j];
        if (atomicAdd(&cnt[w], 1u) == 0u) touched[atomicAdd(&s_nt, 1u)] = w;
      }
    }
    __syncthreads();
    if (c.full5) {
      // Second walk over the same wedges with the final counts: the wedge
      // v-u-w lies on cnt[w] - 1 cycles of the pair (v, w), and each puts
      // its edges vu, uw and its middle u on them (4-cycles per edge and
      // per vertex, weighted by t_e and d_u - 2).
      for (int64_t i = c.rowptr[v] + wv; i < ev; i += nwv) {
        const int32_t u = c.col[i];
        const unsigned long long tvu = c.t_csr[i];
        const unsigned long long du2 = c.deg[u] - 2;
        const int64_t bu = c.rowptr[u];
        const int64_t eu = wave_first_greater(
            c.col, static_cast<int32_t>(v) - 1, bu, c.rowptr[u + 1]);
        for (int64_t j = bu + lane; j < eu; j += 64) {
          const int32_t w = c.col[j];
          // The counts were made by atomics (in L2): read them past L1.
          const unsigned long long x =
              __hip_atomic_load(&cnt[w], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT) - 1u;
          ban += x * du2;
          tq += x * (tvu + c.t_csr[j]);
        }
      }
      __syncthreads();  // every count read before the clear below
    }
    const unsigned nt = s_nt;
    const unsigned long long dv = c.deg[v];
    for (unsigned k = threadIdx.x; k < nt; k += kC4Threads) {
      // Read and clear in one atomic (the counts were made by atomics).
      const int32_t w = touched[k];
      const unsigned long long x = atomicExch(&cnt[w], 0u);
      const unsigned long long cyc = choose2(x);
      acc += cyc;
      if (c.full5) ban += cyc * (dv + c.deg[w] - 4);
    }
    __syncthreads();  // s_v, s_nt and the list are reused
  }
  acc = wave_sum_u64(acc);
  ban = wave_sum_u64(ban);
  tq = wave_sum_u64(tq);
  if (lane == 0) {
    s_acc[wv] = acc;
    s_ban[wv] = ban;
    s_tq[wv] = tq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, b = 0, q = 0;
    for (int k = 0; k < nwv; ++k) {
      t += s_acc[k];
      b += s_ban[k];
      q += s_tq[k];
    }
    if (t) atomicAdd(&c.acc[kCensusC4], t);
    if (b) atomicAdd(&c.acc[kCensusBanner], b);
    if (q) atomicAdd(&c.acc[kCensusCycEdgeT], q);
  }
}

}  // namespace

int census_c4_launch(const CensusArgs &c, hipStream_t s) {
  hipLaunchKernelGGL(census_c4, dim3(std::max(c.c4_groups, 1)),
                     dim3(kC4Threads), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census_waves(int cus) { return cus * 16; }

int64_t census_k5_slab_len(int64_t max_out, int64_t max_deg) {
  const int64_t L = std::max<int64_t>(max_out, 1);
  const int64_t D = std::max<int64_t>(max_deg, 1);
  return ((4 * L + D + 1) & ~int64_t(1)) + 2 * L * ((L + 63) / 64);
}

int census_k5_launch(const CensusArgs &c, hipStream_t s) {
  const int blocks = std::max(c.k5_waves / kWaves, 1);
  hipLaunchKernelGGL(census_k5, dim3(blocks), dim3(kWaves * 64), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census_dag_sources(const int64_t *drp, int32_t n, int32_t *dsrc, int cus,
                       hipStream_t s) {
  const int gv = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>((int64_t(n) + 255) / 256, int64_t(cus) * 16)));
  hipLaunchKernelGGL(dag_sources, dim3(gv), dim3(256), 0, s, drp, n, dsrc);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census_launch(const CensusArgs &c, int cus, hipStream_t s) {
  const int gv = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>((int64_t(c.n) + 255) / 256, int64_t(cus) * 16)));
  const int blocks = census_waves(cus) / kWaves;
  hipLaunchKernelGGL(census_triangles, dim3(blocks), dim3(kWaves * 64), 0, s,
                     c);
  hipLaunchKernelGGL(census_sums, dim3(gv), dim3(256), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

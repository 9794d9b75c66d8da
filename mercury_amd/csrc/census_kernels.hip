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
// A second pass over the DAG arcs sums C(t_e, 2) and C(d_v, 3).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "census.hpp"
#include "wave.hpp"

namespace mgpm {

namespace {

constexpr int kWaves = 4;
constexpr int kPairMax = 16;  // |C| up to this: K4 test by flattened pairs

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
  unsigned long long tri = 0, tdeg = 0, k4 = 0;
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
    for (int p0 = 0; p0 < ls; p0 += 64) {
      const int p = p0 + lane;
      bool hit = false;
      int32_t z = 0;
      if (p < ls) {
        z = S[p];
        const int j = lower_bound_g(L, ll, z);
        if (j < ll && L[j] == z) {
          hit = true;
          atomicAdd(&c.t[sb + p], 1u);
          atomicAdd(&c.t[lb + j], 1u);
          tdeg += static_cast<unsigned long long>(c.deg[z]);
        }
      }
      const unsigned long long mask = __ballot(hit);
      if (hit) slab[cnt + bits_below(mask)] = z;
      cnt += __popcll(mask);
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
      const int np = cnt * cnt;
      for (int p0 = 0; p0 < np; p0 += 64) {
        const int p = p0 + lane;
        if (p >= np) continue;
        const int i = p / cnt, j = p - i * cnt;
        if (i == j) continue;
        const int32_t z = slab[i], q = slab[j];
        const int64_t bz = c.drp[z];
        if (contains_g(c.dcol + bz, static_cast<int>(c.drp[z + 1] - bz), q))
          ++k4;
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
  if (lane == 0) {
    if (tri) atomicAdd(&c.acc[kCensusTriangles], tri);
    if (tdeg) atomicAdd(&c.acc[kCensusTriDegrees], tdeg);
    if (k4) atomicAdd(&c.acc[kCensusK4], k4);
  }
}

// Sums over vertices (C(d,3)) and over DAG arcs (C(t,2) and
// (d_x - t - 1)(d_y - t - 1)).
__global__ void census_sums(CensusArgs c) {
  unsigned long long s3 = 0, s2 = 0, sp = 0;
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < c.n;
       v += int64_t(gridDim.x) * blockDim.x) {
    const unsigned long long d = static_cast<unsigned long long>(c.deg[v]);
    s3 += d >= 3 ? d * (d - 1) * (d - 2) / 6 : 0;
    for (int64_t a = c.drp[v]; a < c.drp[v + 1]; ++a) {
      const unsigned long long t = c.t[a];
      const unsigned long long dy =
          static_cast<unsigned long long>(c.deg[c.dcol[a]]);
      s2 += t * (t - 1) / 2;
      sp += (d - t - 1) * (dy - t - 1);
    }
  }
  s3 = wave_sum_u64(s3);
  s2 = wave_sum_u64(s2);
  sp = wave_sum_u64(sp);
  if (lane64() == 0) {
    if (s3) atomicAdd(&c.acc[kCensusDeg3], s3);
    if (s2) atomicAdd(&c.acc[kCensusEdgePairs], s2);
    if (sp) atomicAdd(&c.acc[kCensusPathPairs], sp);
  }
}

// 4-cycles as subgraphs, each counted once at its highest-ranked vertex v
// (rank: degree, then id -- the DAG order): for every neighbour u of v
// ranked below v and every neighbour w of u ranked below v, count the
// wedge v-u-w at w; a cycle v-u-w-u' is one pair of wedges at w, so v adds
// sum_w C(wedges(w), 2).  One workgroup per v (vertices handed out by an
// atomic cursor), with its own dense counter array over all vertices
// (left zeroed) and the list of the w it touched.
__device__ __forceinline__ bool rank_below(const int32_t *__restrict__ deg,
                                           int32_t a, int32_t b) {
  const int32_t da = deg[a], db = deg[b];
  return da < db || (da == db && a < b);
}

constexpr int kC4Threads = 256;

__global__ __launch_bounds__(kC4Threads) void census_c4(CensusArgs c) {
  uint32_t *cnt = c.wedges + size_t(blockIdx.x) * size_t(c.n);
  int32_t *touched = c.touched + size_t(blockIdx.x) * size_t(c.n);
  __shared__ unsigned s_v, s_nt;
  __shared__ unsigned long long s_acc[kC4Threads / 64];
  const int lane = lane64(), wv = threadIdx.x >> 6;
  constexpr int nwv = kC4Threads / 64;
  unsigned long long acc = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      s_v = atomicAdd(c.cursor + 1, 1u);
      s_nt = 0;
    }
    __syncthreads();
    const unsigned v = s_v;
    if (v >= static_cast<unsigned>(c.n)) break;
    const int64_t ev = c.rowptr[v + 1];
    for (int64_t i = c.rowptr[v] + wv; i < ev; i += nwv) {
      const int32_t u = c.col[i];
      if (!rank_below(c.deg, u, static_cast<int32_t>(v))) continue;
      const int64_t eu = c.rowptr[u + 1];
      for (int64_t j = c.rowptr[u] + lane; j < eu; j += 64) {
        const int32_t w = c.col[j];
        if (!rank_below(c.deg, w, static_cast<int32_t>(v))) continue;
        if (atomicAdd(&cnt[w], 1u) == 0u) touched[atomicAdd(&s_nt, 1u)] = w;
      }
    }
    __syncthreads();
    const unsigned nt = s_nt;
    for (unsigned k = threadIdx.x; k < nt; k += kC4Threads) {
      // Read and clear in one atomic (the counts were made by atomics).
      const unsigned long long x = atomicExch(&cnt[touched[k]], 0u);
      acc += x * (x - 1) / 2;
    }
    __syncthreads();  // s_v, s_nt and the list are reused
  }
  acc = wave_sum_u64(acc);
  if (lane == 0) s_acc[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < nwv; ++k) t += s_acc[k];
    if (t) atomicAdd(&c.acc[kCensusC4], t);
  }
}

}  // namespace

int census_c4_launch(const CensusArgs &c, hipStream_t s) {
  hipLaunchKernelGGL(census_c4, dim3(std::max(c.c4_groups, 1)),
                     dim3(kC4Threads), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int census_waves(int cus) { return cus * 16; }

int census_launch(const CensusArgs &c, int cus, hipStream_t s) {
  const int gv = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>((int64_t(c.n) + 255) / 256, int64_t(cus) * 16)));
  hipLaunchKernelGGL(dag_sources, dim3(gv), dim3(256), 0, s, c.drp, c.n,
                     c.dsrc);
  const int blocks = census_waves(cus) / kWaves;
  hipLaunchKernelGGL(census_triangles, dim3(blocks), dim3(kWaves * 64), 0, s,
                     c);
  hipLaunchKernelGGL(census_sums, dim3(gv), dim3(256), 0, s, c);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mgpm

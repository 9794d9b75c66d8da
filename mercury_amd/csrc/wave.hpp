// Wave64 building blocks for gfx950 (CDNA4): every helper assumes the full
// 64-lane wavefront is active when called.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgpm {

__device__ __forceinline__ int lane64() { return __lane_id(); }

// Cross-lane handoff through LDS inside one wavefront.  The hardware runs a
// wave's LDS instructions in order, but without a fence the compiler may
// forward a lane's own store into its later load (or reorder them), so
// every write-by-one-lane / read-by-another step goes through this.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS-typed int pointer: accesses through it are ds_read/ds_write.
typedef __attribute__((address_space(3))) int32_t lds_i32;

__device__ __forceinline__ lds_i32 *as_lds(int32_t *p) {
  return (lds_i32 *)p;
}

// Compiler-only ordering point for LDS handoffs between the lanes of one
// wave: a wave's DS instructions execute and complete in program order, so
// keeping the compiler from reordering, merging or forwarding them across
// this point is all that is needed (no s_waitcnt, unlike a fence or a
// volatile access).
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// Number of set bits of `m` strictly below the calling lane.
__device__ __forceinline__ int bits_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(
                                       static_cast<uint32_t>(m), 0u));
}

// Inclusive prefix sum across the wave.
__device__ __forceinline__ int wave_incl_scan(int x) {
  const int lane = lane64();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(
    unsigned long long x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

__device__ __forceinline__ int wave_max_i32(int x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int y = __shfl_xor(x, d, 64);
    x = y > x ? y : x;
  }
  return x;
}

__device__ __forceinline__ int wave_sum_i32(int x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Uniform broadcast of lane 0's value (scalar register).
__device__ __forceinline__ uint32_t wave_first(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// The same for a 64-bit value, as two 32-bit halves.  The builtin returns
// int: each half goes back to uint32_t before widening, so a low half of
// 2^31 or more is not sign-extended into the high half.
__device__ __forceinline__ uint64_t wave_first_u64(uint64_t x) {
  const uint32_t lo = wave_first(static_cast<uint32_t>(x));
  const uint32_t hi = wave_first(static_cast<uint32_t>(x >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Multiplicative (Fibonacci) hash of a vertex id into 2^logp slots.
__device__ __forceinline__ uint32_t slot_of(int32_t key, int logp) {
  return (static_cast<uint32_t>(key) * 0x9E3779B1u) >> (32 - logp);
}

// First position in [lo, hi) of a sorted int32 row whose value exceeds
// key (hi if none), found by the whole wave: 64 evenly spaced pivots per
// round narrow the range to one gap, then its last (at most 64) positions
// are tested at once.  Every lane of the wave must call it with the same
// arguments; every lane returns it.
__device__ inline int64_t wave_first_greater(const int32_t *__restrict__ row,
                                             int32_t key, int64_t lo,
                                             int64_t hi) {
  const int lane = lane64();
  while (hi - lo > 64) {
    const int64_t span = hi - lo;  // pivot j at lo + span (j + 1) / 65
    const int64_t p = lo + span * (lane + 1) / 65;
    const unsigned long long up = __ballot(row[p] > key);
    const int f = up ? __builtin_ctzll(up) : 64;  // first pivot above key
    const int64_t nlo = f == 0 ? lo : lo + span * f / 65 + 1;
    if (f < 64) hi = lo + span * (f + 1) / 65;
    lo = nlo;
  }
  const int64_t p = lo + lane;
  const unsigned long long up = __ballot(p < hi && row[p] > key);
  return up ? lo + __builtin_ctzll(up) : hi;
}

// ceil(log2(x)) for x >= 2.
__device__ __forceinline__ int ceil_log2(int x) { return 32 - __clz(x - 1); }

}  // namespace mgpm

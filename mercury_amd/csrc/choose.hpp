// Binomial coefficients C(x, 2), C(x, 3) and C(x, 4) exact modulo 2^64 for
// every 64-bit x: the divisor is taken out of the consecutive factors
// before they are multiplied, so no product wraps before a division (the
// plain x(x-1)(x-2)/6 wraps from x ~ 2.6 million on).  Counts are uint64
// modulo 2^64 throughout, like the counters they feed.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MGPM_HD __host__ __device__ __forceinline__
#else
#define MGPM_HD inline
#endif

namespace mgpm {

MGPM_HD uint64_t choose2(uint64_t x) {
  if (x < 2) return 0;
  uint64_t a = x, b = x - 1;
  if (a % 2 == 0) a /= 2;
  else b /= 2;
  return a * b;
}

MGPM_HD uint64_t choose3(uint64_t x) {
  if (x < 3) return 0;
  uint64_t f[3] = {x, x - 1, x - 2};
  // One of three consecutive integers is a multiple of 3, one is even;
  // dividing by 3 keeps a factor's power of two.
  if (f[0] % 3 == 0) f[0] /= 3;
  else if (f[1] % 3 == 0) f[1] /= 3;
  else f[2] /= 3;
  if (f[0] % 2 == 0) f[0] /= 2;
  else f[1] /= 2;
  return f[0] * f[1] * f[2];
}

MGPM_HD uint64_t choose4(uint64_t x) {
  if (x < 4) return 0;
  uint64_t f[4] = {x, x - 1, x - 2, x - 3};
  int i3 = 0;
  while (f[i3] % 3 != 0) ++i3;  // one of four consecutive integers
  f[i3] /= 3;
  // Of four consecutive integers one is 0 mod 4 and another 2 mod 4.
  int i4 = 0;
  while (f[i4] % 4 != 0) ++i4;
  f[i4] /= 4;
  int i2 = 0;
  while (i2 == i4 || f[i2] % 2 != 0) ++i2;
  f[i2] /= 2;
  return f[0] * f[1] * f[2] * f[3];
}

}  // namespace mgpm

// Integer division and remainder by a literal divisor without a divide
// instruction: Granlund and Montgomery, "Division by Invariant Integers
// using Multiplication" (PLDI 1994), the round-up variant with a 65-bit
// magic number (their Figure 4.1), on |n| for signed operands.
//
// q = trunc(n / d), r = n - q*d (the sign of r follows n), for d != 0.
// i64::MIN / -1 overflows and is reported by the caller, not here.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <cstdint>
#endif

#if !defined(__HIPCC__)
// plain C++ (the host unit test of this header)
#define __host__
#define __device__
#endif

namespace oxr {

struct DivMagic {
  uint64_t m;      // low 64 bits of the 65-bit magic (2^64 + m)
  uint64_t d;      // |divisor|
  int32_t shift;   // l - 1, or log2(d) for a power of two
  int32_t kind;    // 0: d == 1, 1: power of two, 2: general
  int32_t neg;     // divisor < 0
  int32_t pad;
};

inline DivMagic div_magic(int64_t divisor) {
  DivMagic g{};
  g.neg = divisor < 0;
  const uint64_t d = divisor < 0 ? (uint64_t)0 - (uint64_t)divisor : (uint64_t)divisor;
  g.d = d;
  if (d == 1) {
    g.kind = 0;
    return g;
  }
  int l = 0;
  while (l < 64 && ((uint64_t)1 << l) < d) l++;  // l = ceil(log2 d)
  if (((uint64_t)1 << l) == d) {
    g.kind = 1;
    g.shift = l;
    return g;
  }
  // m = floor(2^(64+l) / d) - 2^64 + 1, computed with 128-bit arithmetic
  const unsigned __int128 num = (unsigned __int128)1 << (64 + l);
  const unsigned __int128 q = num / d;
  g.m = (uint64_t)(q - ((unsigned __int128)1 << 64) + 1);
  g.kind = 2;
  g.shift = l - 1;
  return g;
}

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Unsigned quotient of n by g.d.
__host__ __device__ inline uint64_t udiv_magic(uint64_t n, const DivMagic &g) {
  if (g.kind == 0) return n;
  if (g.kind == 1) return n >> g.shift;
  const uint64_t t = mulhi64(g.m, n);
  return (t + ((n - t) >> 1)) >> g.shift;
}

// Signed 64-bit n / d and n % d (truncating), d = the divisor the magic was
// made for. Results wrap for i64::MIN / -1 (the caller reports it).
__host__ __device__ inline void sdivmod_magic(int64_t n, const DivMagic &g, int64_t *q, int64_t *r) {
  const bool nneg = n < 0;
  const uint64_t an = nneg ? (uint64_t)0 - (uint64_t)n : (uint64_t)n;
  const uint64_t uq = udiv_magic(an, g);
  const uint64_t ur = an - uq * g.d;
  const bool qneg = nneg != (bool)g.neg;
  *q = (int64_t)(qneg ? (uint64_t)0 - uq : uq);
  *r = (int64_t)(nneg ? (uint64_t)0 - ur : ur);
}

}  // namespace oxr

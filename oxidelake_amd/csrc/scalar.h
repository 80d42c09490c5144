// Scalar SQL semantics on the device, shared by the expression interpreter
// (expr.h) and the kernels generated per query at run time (jit.hip), so the
// two paths give bit-identical results:
//  - DataFusion's float order for comparisons: -0.0 read as +0.0, then IEEE
//    totalOrder (crates/oxidelake-device/src/cpu/kernels.rs:62-69, 825-830);
//  - wrapping integer arithmetic, truncating division, a zero divisor or
//    MIN / -1 reported as an error, MIN % -1 = 0;
//  - NaN results as x86-64 SSE produces them.
// This header compiles under hipcc and under hipRTC, where no standard
// header exists and the fixed-width integer types live in __hip_internal.
#pragma once

#if defined(__HIPCC_RTC__)
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
using __hip_internal::uint8_t;
#else
#include <cstdint>
#endif

#include "divlit.h"

namespace oxr {

__host__ __device__ inline int64_t order_key64(double x) {
  if (x == 0.0) x = 0.0;
  int64_t b;
  __builtin_memcpy(&b, &x, 8);
  return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}

__host__ __device__ inline int32_t order_key32(float x) {
  if (x == 0.0f) x = 0.0f;
  int32_t b;
  __builtin_memcpy(&b, &x, 4);
  return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

__device__ inline double as_f64(uint64_t v) {
  double d;
  __builtin_memcpy(&d, &v, 8);
  return d;
}
__device__ inline uint64_t f64_bits(double d) {
  uint64_t v;
  __builtin_memcpy(&v, &d, 8);
  return v;
}
__device__ inline float as_f32(uint64_t v) {
  uint32_t u = (uint32_t)v;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ inline uint64_t f32_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (uint64_t)u;
}

constexpr int64_t kI64Min = (int64_t)(~(uint64_t)0 >> 1) * -1 - 1;
constexpr int64_t kI32Min = -2147483647ll - 1;

// Integer division of widened values in type t (32- or 64-bit), truncating;
// a zero divisor, or the overflowing MIN / -1, in a valid row sets *div0.
__device__ inline uint64_t idiv(bool mod, bool i32, uint64_t a, uint64_t b, bool valid, bool *div0) {
  const int64_t x = (int64_t)a, y = (int64_t)b;
  if (y == 0) {
    if (valid) *div0 = true;
    return 0;
  }
  const int64_t lo = i32 ? kI32Min : kI64Min;
  if (x == lo && y == -1) {
    if (!mod && valid) *div0 = true;
    return 0;
  }
  uint64_t z = (uint64_t)(mod ? x % y : x / y);
  return i32 ? (uint64_t)(int64_t)(int32_t)(uint32_t)z : z;
}

// x / lit or x % lit for a non-zero integer literal divisor.
__device__ inline uint64_t idiv_lit(bool mod, bool i32, uint64_t a, const DivMagic &g, bool valid,
                                    bool *div0) {
  const int64_t x = (int64_t)a;
  const int64_t lo = i32 ? kI32Min : kI64Min;
  if (g.neg && g.d == 1 && x == lo) {
    if (!mod && valid) *div0 = true;
    return 0;
  }
  int64_t q, r;
  sdivmod_magic(x, g, &q, &r);
  uint64_t z = (uint64_t)(mod ? r : q);
  return i32 ? (uint64_t)(int64_t)(int32_t)(uint32_t)z : z;
}

__device__ inline uint64_t wrap32(uint64_t z) { return (uint64_t)(int64_t)(int32_t)(uint32_t)z; }

// NaN results as x86-64 SSE produces them (the platform the reference's CPU
// path and its conformance suite run on): a NaN operand propagates, quieted,
// the first operand's first; an invalid operation on non-NaN operands
// (0 * inf, inf - inf, 0 / 0) gives the negative default NaN. The GPU's own
// default NaN is positive, and the sign matters to totalOrder comparisons.
__device__ inline uint64_t x86_nan64(uint64_t z, uint64_t a, uint64_t b) {
  const uint64_t exp = 0x7FF0000000000000ull, man = 0x000FFFFFFFFFFFFFull;
  const auto isnan = [&](uint64_t v) { return (v & exp) == exp && (v & man) != 0; };
  if (!isnan(z)) return z;
  if (isnan(a)) return a | 0x0008000000000000ull;
  if (isnan(b)) return b | 0x0008000000000000ull;
  return 0xFFF8000000000000ull;
}

__device__ inline uint64_t x86_nan32(uint64_t z, uint64_t a, uint64_t b) {
  const uint32_t exp = 0x7F800000u, man = 0x007FFFFFu;
  const auto isnan = [&](uint32_t v) { return (v & exp) == exp && (v & man) != 0; };
  if (!isnan((uint32_t)z)) return z;
  if (isnan((uint32_t)a)) return (uint64_t)((uint32_t)a | 0x00400000u);
  if (isnan((uint32_t)b)) return (uint64_t)((uint32_t)b | 0x00400000u);
  return 0xFFC00000ull;
}

// ------------------------------------------- typed forms for generated code
// Each takes and returns values of the SQL type itself.
#define OXR_F64_OP(name, op)                                                            \
  __device__ inline double name(double a, double b) {                                   \
    return as_f64(x86_nan64(f64_bits(op(a, b)), f64_bits(a), f64_bits(b)));              \
  }
#define OXR_F32_OP(name, op)                                                            \
  __device__ inline float name(float a, float b) {                                      \
    return as_f32(x86_nan32(f32_bits(op(a, b)), f32_bits(a), f32_bits(b)));              \
  }
OXR_F64_OP(f64_add, __dadd_rn)
OXR_F64_OP(f64_sub, __dsub_rn)
OXR_F64_OP(f64_mul, __dmul_rn)
OXR_F64_OP(f64_div, __ddiv_rn)
OXR_F32_OP(f32_add, __fadd_rn)
OXR_F32_OP(f32_sub, __fsub_rn)
OXR_F32_OP(f32_mul, __fmul_rn)
OXR_F32_OP(f32_div, __fdiv_rn)
#undef OXR_F64_OP
#undef OXR_F32_OP
__device__ inline double f64_mod(double a, double b) { return fmod(a, b); }
__device__ inline float f32_mod(float a, float b) { return fmodf(a, b); }

__device__ inline int32_t i32_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ inline int32_t i32_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ inline int32_t i32_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ inline int64_t i64_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
__device__ inline int64_t i64_sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
__device__ inline int64_t i64_mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }

__device__ inline int32_t i32_divmod(bool mod, int32_t a, int32_t b, bool valid, bool *err) {
  return (int32_t)(int64_t)idiv(mod, true, (uint64_t)(int64_t)a, (uint64_t)(int64_t)b, valid, err);
}
__device__ inline int64_t i64_divmod(bool mod, int64_t a, int64_t b, bool valid, bool *err) {
  return (int64_t)idiv(mod, false, (uint64_t)a, (uint64_t)b, valid, err);
}
__device__ inline int32_t i32_divmod_lit(bool mod, int32_t a, const DivMagic &g, bool valid, bool *err) {
  return (int32_t)(int64_t)idiv_lit(mod, true, (uint64_t)(int64_t)a, g, valid, err);
}
__device__ inline int64_t i64_divmod_lit(bool mod, int64_t a, const DivMagic &g, bool valid, bool *err) {
  return (int64_t)idiv_lit(mod, false, (uint64_t)a, g, valid, err);
}

}  // namespace oxr

// Scalar expressions: the host compiler from the ABI's postfix `oxr_expr`
// (include/oxide_rocm.h) to the device program format, and the device
// interpreter that evaluates a program over V consecutive rows per lane.
//
// The reference's plan vocabulary (crates/oxidelake-core/src/params.rs:12-139)
// is `Compare{column, op, literal}` and `And`, and its CPU semantics are
// crates/oxidelake-device/src/cpu/kernels.rs:60-141. This grammar is a
// superset (Or, Int32/Float32, arithmetic, expression operands) with
// DataFusion's numeric coercion.
//
// Device form:
//  - a predicate is a list of leaves (comparisons) and a postfix boolean
//    program over them (LEAF i / AND / OR). For AND/OR trees a NULL leaf can
//    be read as false: Kleene AND/OR then keep exactly the rows the
//    three-valued result keeps (a NULL result drops the row like false).
//  - a value is a three-address program over 4 registers per row.
//  - a leaf `column op literal` (the reference's whole v1 grammar) takes a
//    fast path: one load, one key transform, one integer compare.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"

namespace oxr {

constexpr int kMaxIns = 48;
constexpr int kMaxLeaves = 16;
constexpr int kMaxBool = 32;
constexpr int kMaxCols = 16;
constexpr int kMaxOuts = 8;
constexpr int kRegs = 4;

enum : uint8_t { I_LOAD = 1, I_LIT, I_CAST, I_ADD, I_SUB, I_MUL, I_DIV, I_MOD };
enum : uint8_t { B_LEAF = 0, B_AND = 0x40, B_OR = 0x80 };

struct Ins {
  uint8_t op, type, dst, a;
  uint8_t b, from, pad0, pad1;
  int32_t col, pad2;
  int64_t lit;  // literal bits: int64, double bits, or float bits
};

struct Leaf {
  uint8_t fast, op, ctype, pad;
  int16_t col;
  uint16_t pstart, plen;
  int16_t pad2;
  int64_t key;  // fast path: literal as a compare key of ctype
};

struct DCol {
  const void *values;
  const uint32_t *valid;  // NULL = all valid
  int32_t type, dim;
};

// A compiled value program: [start, start+len) of the instruction array,
// result in register 0.
struct ValueProg {
  int16_t col = -1;  // >= 0: a bare column (no instructions)
  uint16_t start = 0, len = 0;
  int32_t type = 0;
  bool nullable = false;
};

struct Program {
  std::vector<Ins> ins;
  std::vector<Leaf> leaves;
  std::vector<uint8_t> bprog;
  bool has_div = false;
};

// Compiles a predicate (postfix boolean expression) into `p`'s leaves and
// boolean program. Column types come from `cols`.
int compile_predicate(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p);
// Compiles a value expression; appends instructions to p->ins.
int compile_value(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p, ValueProg *out);
// Whether a predicate is in this backend's grammar (no failure recorded).
bool predicate_supported(const oxr_expr *e, std::string *why);

// DataFusion numeric coercion.
inline int32_t coerce(int32_t a, int32_t b) {
  if (a == OXR_FLOAT64 || b == OXR_FLOAT64) return OXR_FLOAT64;
  if (a == OXR_FLOAT32 || b == OXR_FLOAT32) return OXR_FLOAT32;
  if (a == OXR_INT64 || b == OXR_INT64) return OXR_INT64;
  return OXR_INT32;
}

// ---------------------------------------------------------------- device
__host__ __device__ inline int64_t order_key64(double x) {
  // -0.0 read as +0.0, then IEEE totalOrder as a signed integer
  // (kernels.rs:62-69, 825-830).
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

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef uint64_t v2u __attribute__((ext_vector_type(2)));

// Loads V consecutive rows [row0, row0+V) of column c. vmask bit q is set
// when row0+q < n and the row is valid. Values are widened to 64 bits:
// Int32 sign-extended, floats as their bits.
template <int V>
__device__ inline void load_rows(const DCol &c, uint64_t row0, uint64_t n, uint64_t (&x)[V],
                                 uint32_t &vmask) {
  const bool full = row0 + V <= n;
  uint32_t inb = full ? ((1u << V) - 1) : (row0 >= n ? 0u : ((1u << (uint32_t)(n - row0)) - 1));
  uint32_t vm = inb;
  if (c.valid && inb) {
    // V divides 32 and row0 is a multiple of V: the V bits share one word
    uint32_t w = c.valid[row0 >> 5];
    vm &= (w >> (row0 & 31)) & ((1u << V) - 1);
  }
  vmask = vm;
  if (c.type == OXR_INT32 || c.type == OXR_FLOAT32) {
    const int32_t *p = (const int32_t *)c.values + row0;
    if (full && V == 4) {
      const v4i q = *(const v4i *)p;
      int32_t t[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
      for (int k = 0; k < V; k++)
        x[k] = c.type == OXR_INT32 ? (uint64_t)(int64_t)t[k] : (uint64_t)(uint32_t)t[k];
    } else {
#pragma unroll
      for (int k = 0; k < V; k++) {
        int32_t t = (inb >> k) & 1 ? p[k] : 0;
        x[k] = c.type == OXR_INT32 ? (uint64_t)(int64_t)t : (uint64_t)(uint32_t)t;
      }
    }
  } else {
    const uint64_t *p = (const uint64_t *)c.values + row0;
    if (full && (V % 2) == 0) {
#pragma unroll
      for (int k = 0; k < V; k += 2) {
        const v2u q = *(const v2u *)(p + k);
        x[k] = q[0];
        x[k + 1] = q[1];
      }
    } else {
#pragma unroll
      for (int k = 0; k < V; k++) x[k] = (inb >> k) & 1 ? p[k] : 0;
    }
  }
}

// Casts a widened value of type `from` to type `to`.
__device__ inline uint64_t cast_value(uint64_t v, int from, int to) {
  if (from == to) return v;
  switch (to) {
    case OXR_INT32: return (uint64_t)(int64_t)(int32_t)(int64_t)v;
    case OXR_INT64: return v;  // only Int32 (already sign-extended) widens to Int64
    case OXR_FLOAT32:
      if (from == OXR_INT32 || from == OXR_INT64) return f32_bits((float)(int64_t)v);
      return f32_bits((float)as_f64(v));
    case OXR_FLOAT64:
      if (from == OXR_INT32 || from == OXR_INT64) return f64_bits((double)(int64_t)v);
      if (from == OXR_FLOAT32) return f64_bits((double)as_f32(v));
      return v;
  }
  return v;
}

// a op b in type t. Sets *div0 when an integer division by zero happens in
// a valid row.
__device__ inline uint64_t arith(uint8_t op, int t, uint64_t a, uint64_t b, bool valid,
                                 bool *div0) {
  if (t == OXR_FLOAT64) {
    double x = as_f64(a), y = as_f64(b), z = 0;
    switch (op) {
      case I_ADD: z = x + y; break;
      case I_SUB: z = x - y; break;
      case I_MUL: z = __dmul_rn(x, y); break;
      case I_DIV: z = __ddiv_rn(x, y); break;
      case I_MOD: z = fmod(x, y); break;
    }
    return f64_bits(z);
  }
  if (t == OXR_FLOAT32) {
    float x = as_f32(a), y = as_f32(b), z = 0;
    switch (op) {
      case I_ADD: z = x + y; break;
      case I_SUB: z = x - y; break;
      case I_MUL: z = __fmul_rn(x, y); break;
      case I_DIV: z = __fdiv_rn(x, y); break;
      case I_MOD: z = fmodf(x, y); break;
    }
    return f32_bits(z);
  }
  // integers: wrapping + - *, truncating / %; Int32 results wrap to 32 bits
  int64_t x = (int64_t)a, y = (int64_t)b;
  uint64_t z = 0;
  switch (op) {
    case I_ADD: z = a + b; break;
    case I_SUB: z = a - b; break;
    case I_MUL: z = a * b; break;
    case I_DIV:
    case I_MOD: {
      if (y == 0) {
        if (valid) *div0 = true;
        return 0;
      }
      const int64_t lo = t == OXR_INT32 ? (int64_t)INT32_MIN : INT64_MIN;
      if (x == lo && y == -1) {
        if (op == I_DIV && valid) *div0 = true;  // overflow: an error like /0
        return 0;
      }
      z = (uint64_t)(op == I_DIV ? x / y : x % y);
      break;
    }
  }
  if (t == OXR_INT32) z = (uint64_t)(int64_t)(int32_t)(uint32_t)z;
  return z;
}

template <int V>
struct Regs {
  uint64_t v[kRegs][V];
  uint32_t m[kRegs];
};

// Runs instructions [start, start+len) of `ins` over V rows.
template <int V>
__device__ inline void run_program(const Ins *ins, int start, int len, const DCol *cols,
                                   uint64_t row0, uint64_t n, uint64_t (&r)[kRegs][V],
                                   uint32_t (&m)[kRegs], bool *div0) {
  for (int k = start; k < start + len; k++) {
    const Ins in = ins[k];
    uint64_t t[V];
    uint32_t tm = 0;
    if (in.op == I_LOAD) {
      load_rows<V>(cols[in.col], row0, n, t, tm);
    } else if (in.op == I_LIT) {
#pragma unroll
      for (int q = 0; q < V; q++) t[q] = (uint64_t)in.lit;
      tm = (1u << V) - 1;
    } else {
      uint64_t a[V], b[V];
      uint32_t am = 0, bm = 0;
      // uniform register selects
#pragma unroll
      for (int i = 0; i < kRegs; i++) {
        if (in.a == i) {
#pragma unroll
          for (int q = 0; q < V; q++) a[q] = r[i][q];
          am = m[i];
        }
        if (in.b == i) {
#pragma unroll
          for (int q = 0; q < V; q++) b[q] = r[i][q];
          bm = m[i];
        }
      }
      if (in.op == I_CAST) {
#pragma unroll
        for (int q = 0; q < V; q++) t[q] = cast_value(a[q], in.from, in.type);
        tm = am;
      } else {
        tm = am & bm;
#pragma unroll
        for (int q = 0; q < V; q++) t[q] = arith(in.op, in.type, a[q], b[q], (tm >> q) & 1, div0);
      }
    }
#pragma unroll
    for (int i = 0; i < kRegs; i++) {
      if (in.dst == i) {
#pragma unroll
        for (int q = 0; q < V; q++) r[i][q] = t[q];
        m[i] = tm;
      }
    }
  }
}

__device__ inline bool compare_keys(int64_t a, int64_t b, int op) {
  switch (op) {
    case OXR_EQ: return a == b;
    case OXR_LT: return a < b;
    case OXR_LTEQ: return a <= b;
    case OXR_GT: return a > b;
    case OXR_GTEQ: return a >= b;
    default: return a != b;
  }
}

// The compare key of a widened value of type t (floats: DataFusion order).
__device__ inline int64_t compare_key(uint64_t v, int t) {
  if (t == OXR_FLOAT64) return order_key64(as_f64(v));
  if (t == OXR_FLOAT32) return (int64_t)order_key32(as_f32(v));
  return (int64_t)v;
}

// V-bit mask of rows where the leaf is true (NULL or out of range: 0).
template <int V>
__device__ inline uint32_t eval_leaf(const Leaf &lf, const Ins *ins, const DCol *cols,
                                     uint64_t row0, uint64_t n, bool *div0) {
  uint32_t out = 0;
  if (lf.fast) {
    uint64_t x[V];
    uint32_t vm;
    const DCol c = cols[lf.col];
    load_rows<V>(c, row0, n, x, vm);
#pragma unroll
    for (int q = 0; q < V; q++) {
      int64_t k = compare_key(cast_value(x[q], c.type, lf.ctype), lf.ctype);
      out |= (uint32_t)compare_keys(k, lf.key, lf.op) << q;
    }
    return out & vm;
  }
  uint64_t r[kRegs][V];
  uint32_t m[kRegs] = {0, 0, 0, 0};
  run_program<V>(ins, lf.pstart, lf.plen, cols, row0, n, r, m, div0);
#pragma unroll
  for (int q = 0; q < V; q++)
    out |= (uint32_t)compare_keys(compare_key(r[0][q], lf.ctype), compare_key(r[1][q], lf.ctype),
                                  lf.op)
           << q;
  return out & m[0] & m[1];
}

// Evaluates the boolean program: V-bit mask of selected rows.
template <int V>
__device__ inline uint32_t eval_predicate(const Leaf *leaves, const uint8_t *bprog, int nb,
                                          const Ins *ins, const DCol *cols, uint64_t row0,
                                          uint64_t n, bool *div0) {
  const uint32_t all = (1u << V) - 1;
  // rows at or past n are never selected, whatever the leaves say (a leaf
  // comparing two literals is true everywhere)
  const uint32_t inb = row0 + V <= n ? all : (row0 >= n ? 0u : ((1u << (uint32_t)(n - row0)) - 1));
  if (nb == 0) return inb;
  uint64_t st = 0;  // stack of V-bit masks
  for (int k = 0; k < nb; k++) {
    const uint8_t b = bprog[k];
    if (b < B_AND) {
      st = (st << V) | eval_leaf<V>(leaves[b], ins, cols, row0, n, div0);
    } else {
      uint64_t top = st & all;
      st >>= V;
      uint64_t nxt = st & all;
      uint64_t res = b == B_AND ? (nxt & top) : (nxt | top);
      st = (st & ~(uint64_t)all) | res;
    }
  }
  return (uint32_t)(st & all) & inb;
}

}  // namespace oxr

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
//  - every input column an expression reads is a "slot": a kernel loads all
//    slots of its rows at once (16-byte vector loads, every load in flight
//    together) and evaluates from registers;
//  - a predicate is a list of leaves (comparisons) and a postfix boolean
//    program over them (LEAF i / AND / OR). For AND/OR trees a NULL leaf can
//    be read as false: Kleene AND/OR then keep exactly the rows the
//    three-valued result keeps (a NULL result drops the row like false);
//  - a value is a three-address program over 4 registers per row; register
//    and slot operands are uniform, so the interpreter branches on them
//    (a scalar branch) and the per-row code uses fixed registers;
//  - a leaf `column op literal` (the reference's whole v1 grammar) takes a
//    fast path: one key transform and one integer compare;
//  - integer / and % by a literal use a multiply-high (divlit.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "divlit.h"
#include "scalar.h"

namespace oxr {

constexpr int kMaxIns = 48;
constexpr int kMaxLeaves = 16;
constexpr int kMaxBool = 32;
constexpr int kMaxCols = 16;
constexpr int kMaxOuts = 8;
constexpr int kMaxSlots = 8;
constexpr int kMaxMagic = 4;
constexpr int kRegs = 4;

// Opcodes: the host fuses each operation with its operand type.
enum : uint8_t {
  I_LOAD = 1,
  I_LIT,
  I_CAST_I_F32,    // Int32/Int64 -> Float32
  I_CAST_I_F64,    // Int32/Int64 -> Float64
  I_CAST_F32_F64,  // Float32 -> Float64
  I_ADD_I32, I_SUB_I32, I_MUL_I32, I_DIV_I32, I_MOD_I32,
  I_ADD_I64, I_SUB_I64, I_MUL_I64, I_DIV_I64, I_MOD_I64,
  I_ADD_F32, I_SUB_F32, I_MUL_F32, I_DIV_F32, I_MOD_F32,
  I_ADD_F64, I_SUB_F64, I_MUL_F64, I_DIV_F64, I_MOD_F64,
  I_DIVL_I32, I_MODL_I32, I_DIVL_I64, I_MODL_I64  // by a literal (magic number)
};
enum : uint8_t { B_LEAF = 0, B_AND = 0x40, B_OR = 0x80 };

struct Ins {
  uint8_t op, type, dst, a;  // a: slot for I_LOAD
  uint8_t pad0, pad1, pad2_, pad3;
  int32_t col, pad2;          // I_LOAD: input column (host side; slot in `a`)
  int64_t lit;                // literal bits, or the magic index for I_DIVL/I_MODL
};

struct Leaf {
  uint8_t fast, op, ctype, slot;
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
  int16_t slot = -1; // its slot, for scalar columns
  uint16_t start = 0, len = 0;
  int32_t type = 0;
  bool nullable = false;
};

struct Program {
  std::vector<Ins> ins;
  std::vector<Leaf> leaves;
  std::vector<uint8_t> bprog;
  std::vector<DivMagic> magic;
  std::vector<int32_t> slot_cols;  // slot -> input column
  bool has_div = false;
  // slot of column c (assigned on first use); -1 when the slots are full
  int slot_of(int32_t c) {
    for (size_t i = 0; i < slot_cols.size(); i++)
      if (slot_cols[i] == c) return (int)i;
    if ((int)slot_cols.size() >= kMaxSlots) return -1;
    slot_cols.push_back(c);
    return (int)slot_cols.size() - 1;
  }
};

// Compiles a predicate (postfix boolean expression) into `p`'s leaves and
// boolean program. Column types come from `cols`.
int compile_predicate(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p);
// Compiles a value expression; appends instructions to p->ins.
int compile_value(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p, ValueProg *out);
// Where the right operand of the binary node ending at end-1 begins (nodes
// [begin, end) form one subtree).
int split_operands(const oxr_expr *e, int begin, int end);
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
typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef uint64_t v2u __attribute__((ext_vector_type(2)));

__device__ inline uint32_t in_bounds(uint64_t row0, uint64_t n, int v) {
  const uint32_t all = (1u << v) - 1;
  return row0 + v <= n ? all : (row0 >= n ? 0u : ((1u << (uint32_t)(n - row0)) - 1));
}

// Loads V consecutive rows [row0, row0+V) of column c. vmask bit q is set
// when row0+q < n and the row is valid. Values are widened to 64 bits:
// Int32 sign-extended, floats as their bits.
template <int V>
__device__ inline void load_rows(const DCol &c, uint64_t row0, uint64_t n, uint64_t (&x)[V],
                                 uint32_t &vmask) {
  const bool full = row0 + V <= n;
  const uint32_t inb = in_bounds(row0, n, V);
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

// The loaded slots of V rows.
template <int V, int NS>
struct Slots {
  uint64_t v[NS][V];
  uint32_t m[NS];
};

template <int V, int NS>
__device__ inline void load_slots(const DCol *cols, const int8_t *slot_cols, int nslots,
                                  uint64_t row0, uint64_t n, Slots<V, NS> &s) {
#pragma unroll
  for (int k = 0; k < NS; k++) {
    if (k < nslots) {
      load_rows<V>(cols[slot_cols[k]], row0, n, s.v[k], s.m[k]);
    } else {
#pragma unroll
      for (int q = 0; q < V; q++) s.v[k][q] = 0;
      s.m[k] = 0;
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

template <int V>
struct Regs {
  uint64_t v[kRegs][V];
  uint32_t m[kRegs];
};

// Applies a fused binary opcode to r[D] and r[D+1] (D static).
template <int V, int D>
__device__ inline void bin_at(Regs<V> &R, const Ins &in, const DivMagic *magic, bool *div0) {
  if constexpr (D + 1 < kRegs) {
    uint64_t(&a)[V] = R.v[D];
    const uint64_t(&b)[V] = R.v[D + 1];
    const uint32_t m = R.m[D] & R.m[D + 1];
#define OXR_ROWS(expr) \
  _Pragma("unroll") for (int q = 0; q < V; q++) a[q] = (expr);
    switch (in.op) {
      case I_ADD_I32: OXR_ROWS(wrap32(a[q] + b[q])) break;
      case I_SUB_I32: OXR_ROWS(wrap32(a[q] - b[q])) break;
      case I_MUL_I32: OXR_ROWS(wrap32(a[q] * b[q])) break;
      case I_DIV_I32: OXR_ROWS(idiv(false, true, a[q], b[q], (m >> q) & 1, div0)) break;
      case I_MOD_I32: OXR_ROWS(idiv(true, true, a[q], b[q], (m >> q) & 1, div0)) break;
      case I_ADD_I64: OXR_ROWS(a[q] + b[q]) break;
      case I_SUB_I64: OXR_ROWS(a[q] - b[q]) break;
      case I_MUL_I64: OXR_ROWS(a[q] * b[q]) break;
      case I_DIV_I64: OXR_ROWS(idiv(false, false, a[q], b[q], (m >> q) & 1, div0)) break;
      case I_MOD_I64: OXR_ROWS(idiv(true, false, a[q], b[q], (m >> q) & 1, div0)) break;
      case I_ADD_F32: OXR_ROWS(x86_nan32(f32_bits(__fadd_rn(as_f32(a[q]), as_f32(b[q]))), a[q], b[q])) break;
      case I_SUB_F32: OXR_ROWS(x86_nan32(f32_bits(__fsub_rn(as_f32(a[q]), as_f32(b[q]))), a[q], b[q])) break;
      case I_MUL_F32: OXR_ROWS(x86_nan32(f32_bits(__fmul_rn(as_f32(a[q]), as_f32(b[q]))), a[q], b[q])) break;
      case I_DIV_F32: OXR_ROWS(x86_nan32(f32_bits(__fdiv_rn(as_f32(a[q]), as_f32(b[q]))), a[q], b[q])) break;
      case I_MOD_F32: OXR_ROWS(f32_bits(fmodf(as_f32(a[q]), as_f32(b[q])))) break;
      case I_ADD_F64: OXR_ROWS(x86_nan64(f64_bits(__dadd_rn(as_f64(a[q]), as_f64(b[q]))), a[q], b[q])) break;
      case I_SUB_F64: OXR_ROWS(x86_nan64(f64_bits(__dsub_rn(as_f64(a[q]), as_f64(b[q]))), a[q], b[q])) break;
      case I_MUL_F64: OXR_ROWS(x86_nan64(f64_bits(__dmul_rn(as_f64(a[q]), as_f64(b[q]))), a[q], b[q])) break;
      case I_DIV_F64: OXR_ROWS(x86_nan64(f64_bits(__ddiv_rn(as_f64(a[q]), as_f64(b[q]))), a[q], b[q])) break;
      case I_MOD_F64: OXR_ROWS(f64_bits(fmod(as_f64(a[q]), as_f64(b[q])))) break;
      case I_DIVL_I32:
      case I_MODL_I32:
      case I_DIVL_I64:
      case I_MODL_I64: {
        const DivMagic g = magic[in.lit];
        const bool mod = in.op == I_MODL_I32 || in.op == I_MODL_I64;
        const bool i32 = in.op == I_DIVL_I32 || in.op == I_MODL_I32;
        OXR_ROWS(idiv_lit(mod, i32, a[q], g, (R.m[D] >> q) & 1, div0))
        return;  // one operand: the mask stays r[D]'s
      }
    }
#undef OXR_ROWS
    R.m[D] = m;
  }
}

template <int V, int D, int NS>
__device__ inline void load_at(Regs<V> &R, const Ins &in, const Slots<V, NS> &S) {
#pragma unroll
  for (int k = 0; k < NS; k++) {
    if (in.a == k) {
#pragma unroll
      for (int q = 0; q < V; q++) R.v[D][q] = S.v[k][q];
      R.m[D] = S.m[k];
    }
  }
}

template <int V, int D>
__device__ inline void lit_at(Regs<V> &R, const Ins &in) {
#pragma unroll
  for (int q = 0; q < V; q++) R.v[D][q] = (uint64_t)in.lit;
  R.m[D] = (1u << V) - 1;
}

template <int V, int D>
__device__ inline void cast_at(Regs<V> &R, const Ins &in) {
  uint64_t(&a)[V] = R.v[D];
  switch (in.op) {
    case I_CAST_I_F32:
#pragma unroll
      for (int q = 0; q < V; q++) a[q] = f32_bits((float)(int64_t)a[q]);
      break;
    case I_CAST_I_F64:
#pragma unroll
      for (int q = 0; q < V; q++) a[q] = f64_bits((double)(int64_t)a[q]);
      break;
    default:  // I_CAST_F32_F64
#pragma unroll
      for (int q = 0; q < V; q++) a[q] = f64_bits((double)as_f32(a[q]));
      break;
  }
}

template <int V, int D, int NS>
__device__ inline void step_at(Regs<V> &R, const Ins &in, const Slots<V, NS> &S,
                               const DivMagic *magic, bool *div0) {
  if (in.op == I_LOAD) load_at<V, D, NS>(R, in, S);
  else if (in.op == I_LIT) lit_at<V, D>(R, in);
  else if (in.op <= I_CAST_F32_F64) cast_at<V, D>(R, in);
  else bin_at<V, D>(R, in, magic, div0);
}

// Runs instructions [start, start+len). The destination register is a
// uniform field: one scalar branch picks code with fixed registers, and the
// fused opcode picks one typed straight-line row loop.
template <int V, int NS>
__device__ inline void run_program(const Ins *ins, int start, int len, const Slots<V, NS> &S,
                                   const DivMagic *magic, Regs<V> &R, bool *div0) {
  for (int k = start; k < start + len; k++) {
    const Ins in = ins[k];
    switch (in.dst) {
      case 0: step_at<V, 0, NS>(R, in, S, magic, div0); break;
      case 1: step_at<V, 1, NS>(R, in, S, magic, div0); break;
      case 2: step_at<V, 2, NS>(R, in, S, magic, div0); break;
      default: step_at<V, 3, NS>(R, in, S, magic, div0); break;
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

template <int V, int NS>
__device__ inline uint32_t fast_leaf(const Leaf &lf, const Slots<V, NS> &S, int from_type) {
  uint32_t out = 0, vm = 0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    if (lf.slot == k) {
      vm = S.m[k];
#pragma unroll
      for (int q = 0; q < V; q++) {
        const int64_t key = compare_key(cast_value(S.v[k][q], from_type, lf.ctype), lf.ctype);
        out |= (uint32_t)compare_keys(key, lf.key, lf.op) << q;
      }
    }
  }
  return out & vm;
}

// V-bit mask of rows where the leaf is true (NULL or out of range: 0).
template <int V, int NS>
__device__ inline uint32_t eval_leaf(const Leaf &lf, const Ins *ins, const DCol *cols,
                                     const Slots<V, NS> &S, const DivMagic *magic, bool *div0) {
  if (lf.fast) return fast_leaf<V, NS>(lf, S, cols[lf.col].type);
  Regs<V> R;
#pragma unroll
  for (int i = 0; i < kRegs; i++) R.m[i] = 0;
  run_program<V, NS>(ins, lf.pstart, lf.plen, S, magic, R, div0);
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < V; q++)
    out |= (uint32_t)compare_keys(compare_key(R.v[0][q], lf.ctype), compare_key(R.v[1][q], lf.ctype),
                                  lf.op)
           << q;
  return out & R.m[0] & R.m[1];
}

// Evaluates the boolean program: V-bit mask of selected rows.
template <int V, int NS>
__device__ inline uint32_t eval_predicate(const Leaf *leaves, const uint8_t *bprog, int nb,
                                          const Ins *ins, const DCol *cols, const Slots<V, NS> &S,
                                          const DivMagic *magic, uint64_t row0, uint64_t n,
                                          bool *div0) {
  const uint32_t all = (1u << V) - 1;
  // rows at or past n are never selected, whatever the leaves say (a leaf
  // comparing two literals is true everywhere)
  const uint32_t inb = in_bounds(row0, n, V);
  if (nb == 0) return inb;
  uint64_t st = 0;  // stack of V-bit masks
  for (int k = 0; k < nb; k++) {
    const uint8_t b = bprog[k];
    if (b < B_AND) {
      st = (st << V) | eval_leaf<V, NS>(leaves[b], ins, cols, S, magic, div0);
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

// The evaluated value of a compiled ValueProg for V rows.
template <int V, int NS>
__device__ inline void eval_value(const ValueProg &vp, const Ins *ins, const Slots<V, NS> &S,
                                  const DivMagic *magic, uint64_t (&x)[V], uint32_t &m, bool *div0) {
  if (vp.slot >= 0) {
#pragma unroll
    for (int k = 0; k < NS; k++) {
      if (vp.slot == k) {
#pragma unroll
        for (int q = 0; q < V; q++) x[q] = S.v[k][q];
        m = S.m[k];
      }
    }
    return;
  }
  Regs<V> R;
#pragma unroll
  for (int i = 0; i < kRegs; i++) R.m[i] = 0;
  run_program<V, NS>(ins, vp.start, vp.len, S, magic, R, div0);
#pragma unroll
  for (int q = 0; q < V; q++) x[q] = R.v[0][q];
  m = R.m[0];
}

// Smallest instantiated slot count holding `ns` slots.
inline int slot_bucket(int ns) { return ns <= 2 ? 2 : ns <= 4 ? 4 : 8; }

}  // namespace oxr

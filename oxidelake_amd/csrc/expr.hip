// Host compiler: ABI postfix expressions -> device programs (see expr.h).
#include <cstring>

#include "expr.h"

namespace oxr {
namespace {

// A postfix subtree: nodes [begin, end) of the expression, its result type
// (0 = boolean) and the stack depth it needs.
struct Sub {
  int begin, end;
  int32_t type;
  int depth;
  bool nullable;
};

int flip(int op) {
  switch (op) {
    case OXR_LT: return OXR_GT;
    case OXR_LTEQ: return OXR_GTEQ;
    case OXR_GT: return OXR_LT;
    case OXR_GTEQ: return OXR_LTEQ;
  }
  return op;
}

bool is_int(int32_t t) { return t == OXR_INT32 || t == OXR_INT64; }

// Parses the postfix expression into subtrees; checks arity and types.
int parse(const oxr_expr *e, const std::vector<DevCol> &cols, std::vector<Sub> *subs) {
  std::vector<Sub> st;
  for (int k = 0; k < e->count; k++) {
    const oxr_expr_node &nd = e->nodes[k];
    switch (nd.kind) {
      case OXR_X_COLUMN: {
        if (nd.column < 0 || nd.column >= (int)cols.size())
          return fail(OXR_EPLAN, "rocm.expr",
                      "column index " + std::to_string(nd.column) + " out of range for a batch with " +
                          std::to_string(cols.size()) + " columns");
        const DevCol &c = cols[nd.column];
        st.push_back({k, k + 1, c.type, 1, c.validity != nullptr});
        break;
      }
      case OXR_X_LITERAL:
        if (nd.type < OXR_INT32 || nd.type > OXR_FLOAT64)
          return fail(OXR_EUNSUPPORTED, "rocm.expr", "literal of an unsupported type");
        st.push_back({k, k + 1, nd.type, 1, false});
        break;
      case OXR_X_ADD:
      case OXR_X_SUB:
      case OXR_X_MUL:
      case OXR_X_DIV:
      case OXR_X_MOD:
      case OXR_X_CMP:
      case OXR_X_AND:
      case OXR_X_OR: {
        if (st.size() < 2) return fail(OXR_EPLAN, "rocm.expr", "malformed postfix expression");
        Sub b = st.back();
        st.pop_back();
        Sub a = st.back();
        st.pop_back();
        Sub r{a.begin, k + 1, 0, 0, a.nullable || b.nullable};
        if (nd.kind == OXR_X_AND || nd.kind == OXR_X_OR) {
          if (a.type != 0 || b.type != 0)
            return fail(OXR_EPLAN, "rocm.expr", "AND/OR of a non-boolean operand");
          r.type = 0;
        } else {
          if (a.type == 0 || b.type == 0)
            return fail(OXR_EPLAN, "rocm.expr", "arithmetic or comparison on a boolean");
          if (a.type == OXR_F32_VECTOR || b.type == OXR_F32_VECTOR)
            return fail(OXR_EUNSUPPORTED, "rocm.expr",
                        "vector columns are only projected as they are");
          if (nd.kind == OXR_X_CMP) {
            if (nd.op < OXR_EQ || nd.op > OXR_NOTEQ)
              return fail(OXR_EPLAN, "rocm.expr", "unknown comparison operator");
            r.type = 0;
          } else {
            r.type = coerce(a.type, b.type);
          }
          r.depth = std::max(a.depth, b.depth + 1);
        }
        st.push_back(r);
        break;
      }
      default:
        return fail(OXR_EPLAN, "rocm.expr", "unknown expression node kind " + std::to_string(nd.kind));
    }
  }
  if (st.size() != 1) return fail(OXR_EPLAN, "rocm.expr", "malformed postfix expression");
  subs->swap(st);
  return OXR_OK;
}

// Splits nodes [begin, end) (a value subtree) into its two operand subtrees.
void split(const oxr_expr *e, int begin, int end, int *mid) {
  // walk back from end-2 to find where the right operand starts
  int need = 1;
  int k = end - 2;
  for (; k >= begin; k--) {
    const int kind = e->nodes[k].kind;
    if (kind == OXR_X_COLUMN || kind == OXR_X_LITERAL) need -= 1;
    else need += 1;
    if (need == 0) break;
  }
  *mid = k;  // right operand = [k, end-1)
}

struct Emitter {
  const oxr_expr *e;
  const std::vector<DevCol> &cols;
  Program *p;

  // Result type of nodes [begin, end) (a value subtree), without emitting.
  int32_t type_of(int begin, int end) {
    const oxr_expr_node &nd = e->nodes[end - 1];
    if (nd.kind == OXR_X_COLUMN) return cols[nd.column].type;
    if (nd.kind == OXR_X_LITERAL) return nd.type;
    if (nd.kind == OXR_X_CMP || nd.kind == OXR_X_AND || nd.kind == OXR_X_OR) return 0;
    int mid;
    split(e, begin, end, &mid);
    return coerce(type_of(begin, mid), type_of(mid, end - 1));
  }

  // Emits code computing nodes [begin, end) into register `reg`; returns
  // its type, or 0 on failure (error recorded).
  int32_t emit(int begin, int end, int reg, bool *nullable) {
    if (reg >= kRegs) {
      fail(OXR_EUNSUPPORTED, "rocm.expr", "expression needs more than 4 registers");
      return 0;
    }
    if ((int)p->ins.size() >= kMaxIns) {
      fail(OXR_EUNSUPPORTED, "rocm.expr", "expression has too many instructions");
      return 0;
    }
    const oxr_expr_node &nd = e->nodes[end - 1];
    Ins in;
    std::memset(&in, 0, sizeof in);
    if (nd.kind == OXR_X_COLUMN) {
      const int slot = p->slot_of(nd.column);
      if (slot < 0) {
        fail(OXR_EUNSUPPORTED, "rocm.expr",
             "an operator reads at most " + std::to_string(kMaxSlots) + " distinct columns");
        return 0;
      }
      in.op = I_LOAD;
      in.col = nd.column;
      in.a = (uint8_t)slot;
      in.type = (uint8_t)cols[nd.column].type;
      in.dst = (uint8_t)reg;
      p->ins.push_back(in);
      *nullable = *nullable || cols[nd.column].validity != nullptr;
      return in.type;
    }
    if (nd.kind == OXR_X_LITERAL) {
      in.op = I_LIT;
      in.type = (uint8_t)nd.type;
      in.dst = (uint8_t)reg;
      if (nd.type == OXR_INT32) in.lit = (int64_t)(int32_t)nd.lit_i;
      else if (nd.type == OXR_INT64) in.lit = nd.lit_i;
      else if (nd.type == OXR_FLOAT32) {
        float f = (float)nd.lit_f;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        in.lit = (int64_t)u;
      } else {
        std::memcpy(&in.lit, &nd.lit_f, 8);
      }
      p->ins.push_back(in);
      return nd.type;
    }
    int mid;
    split(e, begin, end, &mid);
    // integer / or % by a non-zero integer literal: multiply-high by a magic
    // number instead of a division
    const oxr_expr_node &rhs = e->nodes[mid];
    const int64_t dlit = rhs.type == OXR_INT32 ? (int64_t)(int32_t)rhs.lit_i : rhs.lit_i;
    if ((nd.kind == OXR_X_DIV || nd.kind == OXR_X_MOD) && end - 1 - mid == 1 &&
        rhs.kind == OXR_X_LITERAL && is_int(rhs.type) && dlit != 0 && is_int(type_of(begin, mid)) &&
        (int)p->magic.size() < kMaxMagic) {
      const int64_t d = dlit;
      int32_t ta = emit(begin, mid, reg, nullable);
      if (!ta) return 0;
      {
        const int32_t t = coerce(ta, rhs.type);
        if (!cast(reg, ta, t)) return 0;
        Ins dl;
        std::memset(&dl, 0, sizeof dl);
        const bool i32 = t == OXR_INT32;
        dl.op = nd.kind == OXR_X_DIV ? (i32 ? I_DIVL_I32 : I_DIVL_I64) : (i32 ? I_MODL_I32 : I_MODL_I64);
        dl.type = (uint8_t)t;
        dl.dst = (uint8_t)reg;
        dl.a = (uint8_t)reg;
        dl.lit = (int64_t)p->magic.size();
        p->magic.push_back(div_magic(d));
        if ((int)p->ins.size() >= kMaxIns) {
          fail(OXR_EUNSUPPORTED, "rocm.expr", "expression has too many instructions");
          return 0;
        }
        p->ins.push_back(dl);
        return t;
      }
    }
    int32_t ta = emit(begin, mid, reg, nullable);
    if (!ta) return 0;
    int32_t tb = emit(mid, end - 1, reg + 1, nullable);
    if (!tb) return 0;
    const int32_t t = coerce(ta, tb);
    if (!cast(reg, ta, t) || !cast(reg + 1, tb, t)) return 0;
    std::memset(&in, 0, sizeof in);
    int k;
    switch (nd.kind) {
      case OXR_X_ADD: k = 0; break;
      case OXR_X_SUB: k = 1; break;
      case OXR_X_MUL: k = 2; break;
      case OXR_X_DIV: k = 3; break;
      case OXR_X_MOD: k = 4; break;
      default:
        fail(OXR_EPLAN, "rocm.expr", "a boolean inside a value expression");
        return 0;
    }
    // fused opcode: the operation within its type's block of five
    const int base = t == OXR_INT32 ? I_ADD_I32 : t == OXR_INT64 ? I_ADD_I64
                     : t == OXR_FLOAT32 ? I_ADD_F32 : I_ADD_F64;
    in.op = (uint8_t)(base + k);
    if (k >= 3 && is_int(t)) p->has_div = true;
    in.type = (uint8_t)t;
    in.dst = (uint8_t)reg;
    if ((int)p->ins.size() >= kMaxIns) {
      fail(OXR_EUNSUPPORTED, "rocm.expr", "expression has too many instructions");
      return 0;
    }
    p->ins.push_back(in);
    return t;
  }

  // Widens register `reg` from `from` to `to` (coercion never narrows).
  // Int32 values are held sign-extended, so Int32 -> Int64 is free.
  bool cast(int reg, int32_t from, int32_t to) {
    if (from == to || (from == OXR_INT32 && to == OXR_INT64)) return true;
    if ((int)p->ins.size() >= kMaxIns) {
      fail(OXR_EUNSUPPORTED, "rocm.expr", "expression has too many instructions");
      return false;
    }
    Ins in;
    std::memset(&in, 0, sizeof in);
    if (to == OXR_FLOAT32 && is_int(from)) in.op = I_CAST_I_F32;
    else if (to == OXR_FLOAT64 && is_int(from)) in.op = I_CAST_I_F64;
    else if (to == OXR_FLOAT64 && from == OXR_FLOAT32) in.op = I_CAST_F32_F64;
    else {
      fail(OXR_EPLAN, "rocm.expr", "unexpected narrowing cast");
      return false;
    }
    in.type = (uint8_t)to;
    in.dst = (uint8_t)reg;
    p->ins.push_back(in);
    return true;
  }
};

int64_t literal_key(const oxr_expr_node &nd, int32_t ctype) {
  if (ctype == OXR_FLOAT64) {
    double d = nd.type == OXR_FLOAT64   ? nd.lit_f
               : nd.type == OXR_FLOAT32 ? (double)(float)nd.lit_f
               : nd.type == OXR_INT32   ? (double)(int32_t)nd.lit_i
                                        : (double)nd.lit_i;
    return order_key64(d);
  }
  if (ctype == OXR_FLOAT32) {
    float f = nd.type == OXR_FLOAT32 ? (float)nd.lit_f
              : nd.type == OXR_INT32 ? (float)(int32_t)nd.lit_i
                                     : (float)nd.lit_i;
    return (int64_t)order_key32(f);
  }
  if (nd.type == OXR_INT32) return (int64_t)(int32_t)nd.lit_i;
  return nd.lit_i;
}

// Compiles one comparison subtree [begin, end) into a leaf.
int compile_leaf(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p, int begin,
                 int end, const std::vector<Sub> &) {
  if ((int)p->leaves.size() >= kMaxLeaves)
    return fail(OXR_EUNSUPPORTED, "rocm.filter", "predicate has more than 16 comparisons");
  const oxr_expr_node &cmp = e->nodes[end - 1];
  int mid;
  split(e, begin, end, &mid);
  Leaf lf;
  std::memset(&lf, 0, sizeof lf);
  lf.op = (uint8_t)cmp.op;
  const oxr_expr_node &l = e->nodes[begin];
  const oxr_expr_node &r = e->nodes[mid];
  const bool lcol = mid - begin == 1 && l.kind == OXR_X_COLUMN;
  const bool rcol = end - 1 - mid == 1 && r.kind == OXR_X_COLUMN;
  const bool llit = mid - begin == 1 && l.kind == OXR_X_LITERAL;
  const bool rlit = end - 1 - mid == 1 && r.kind == OXR_X_LITERAL;
  if ((lcol && rlit) || (llit && rcol)) {
    const oxr_expr_node &c = lcol ? l : r;
    const oxr_expr_node &lit = lcol ? r : l;
    const int32_t ct = cols[c.column].type;
    if (ct == OXR_F32_VECTOR)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "comparison on a vector column");
    const int slot = p->slot_of(c.column);
    if (slot < 0)
      return fail(OXR_EUNSUPPORTED, "rocm.expr",
                  "an operator reads at most " + std::to_string(kMaxSlots) + " distinct columns");
    lf.fast = 1;
    lf.slot = (uint8_t)slot;
    lf.col = (int16_t)c.column;
    lf.ctype = (uint8_t)coerce(ct, lit.type);
    lf.key = literal_key(lit, lf.ctype);
    if (!lcol) lf.op = (uint8_t)flip(cmp.op);
    p->leaves.push_back(lf);
    return OXR_OK;
  }
  Emitter em{e, cols, p};
  bool nullable = false;
  lf.pstart = (uint16_t)p->ins.size();
  int32_t ta = em.emit(begin, mid, 0, &nullable);
  if (!ta) return OXR_EUNSUPPORTED;
  int32_t tb = em.emit(mid, end - 1, 1, &nullable);
  if (!tb) return OXR_EUNSUPPORTED;
  const int32_t ct = coerce(ta, tb);
  if (!em.cast(0, ta, ct) || !em.cast(1, tb, ct)) return OXR_EUNSUPPORTED;
  lf.plen = (uint16_t)(p->ins.size() - lf.pstart);
  lf.ctype = (uint8_t)ct;
  p->leaves.push_back(lf);
  return OXR_OK;
}

// Walks a boolean subtree [begin, end) emitting leaves and the boolean program.
int compile_bool(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p, int begin,
                 int end, const std::vector<Sub> &subs) {
  const oxr_expr_node &nd = e->nodes[end - 1];
  if (nd.kind == OXR_X_CMP) {
    OXR_TRY(compile_leaf(e, cols, p, begin, end, subs));
    if ((int)p->bprog.size() >= kMaxBool)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "predicate is too long");
    p->bprog.push_back((uint8_t)(B_LEAF + p->leaves.size() - 1));
    return OXR_OK;
  }
  if (nd.kind == OXR_X_AND || nd.kind == OXR_X_OR) {
    int mid;
    split(e, begin, end, &mid);
    OXR_TRY(compile_bool(e, cols, p, begin, mid, subs));
    OXR_TRY(compile_bool(e, cols, p, mid, end - 1, subs));
    if ((int)p->bprog.size() >= kMaxBool)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "predicate is too long");
    p->bprog.push_back(nd.kind == OXR_X_AND ? B_AND : B_OR);
    return OXR_OK;
  }
  return fail(OXR_EPLAN, "rocm.filter", "predicate is not a boolean expression");
}

}  // namespace

int split_operands(const oxr_expr *e, int begin, int end) {
  int mid;
  split(e, begin, end, &mid);
  return mid;
}

int compile_predicate(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p) {
  if (!e || e->count == 0) return OXR_OK;
  if (!e->nodes) return fail(OXR_EPLAN, "rocm.filter", "predicate without nodes");
  std::vector<Sub> subs;
  OXR_TRY(parse(e, cols, &subs));
  if (subs[0].type != 0) return fail(OXR_EPLAN, "rocm.filter", "predicate is not boolean");
  // the boolean stack holds 64/V masks
  return compile_bool(e, cols, p, 0, e->count, subs);
}

int compile_value(const oxr_expr *e, const std::vector<DevCol> &cols, Program *p, ValueProg *out) {
  if (!e || e->count == 0 || !e->nodes) return fail(OXR_EPLAN, "rocm.expr", "empty expression");
  std::vector<Sub> subs;
  OXR_TRY(parse(e, cols, &subs));
  if (subs[0].type == 0)
    return fail(OXR_EUNSUPPORTED, "rocm.expr", "boolean-valued projections are not supported");
  ValueProg v;
  if (e->count == 1 && e->nodes[0].kind == OXR_X_COLUMN) {
    v.col = (int16_t)e->nodes[0].column;
    v.type = cols[v.col].type;
    v.nullable = cols[v.col].validity != nullptr;
    if (v.type != OXR_F32_VECTOR) {
      const int slot = p->slot_of(v.col);
      if (slot < 0)
        return fail(OXR_EUNSUPPORTED, "rocm.expr",
                    "an operator reads at most " + std::to_string(kMaxSlots) + " distinct columns");
      v.slot = (int16_t)slot;
    }
    *out = v;
    return OXR_OK;
  }
  Emitter em{e, cols, p};
  v.start = (uint16_t)p->ins.size();
  bool nullable = false;
  int32_t t = em.emit(0, e->count, 0, &nullable);
  if (!t) return OXR_EUNSUPPORTED;
  v.len = (uint16_t)(p->ins.size() - v.start);
  v.type = t;
  v.nullable = nullable;
  *out = v;
  return OXR_OK;
}

bool predicate_supported(const oxr_expr *e, std::string *why) {
  if (!e || e->count == 0) return true;
  // column types are unknown here: check the shape only, with every
  // referenced column taken as Int64
  int maxcol = -1;
  for (int k = 0; k < e->count; k++)
    if (e->nodes[k].kind == OXR_X_COLUMN) maxcol = std::max(maxcol, e->nodes[k].column);
  std::vector<DevCol> cols(maxcol + 1);
  for (auto &c : cols) c.type = OXR_INT64;
  Program p;
  int rc = compile_predicate(e, cols, &p);
  if (rc != OXR_OK && why) *why = oxr_last_error();
  return rc == OXR_OK;
}

}  // namespace oxr

extern "C" int oxr_supports_predicate(const oxr_expr *predicate) {
  std::string why;
  return oxr::predicate_supported(predicate, &why) ? 1 : 0;
}

// The device runs every hash join and aggregation the C ABI accepts; what it
// declines per call (a key type, a size) comes back as OXR_EUNSUPPORTED from
// that call.
extern "C" int oxr_supports_hash_join(void) { return 1; }
extern "C" int oxr_supports_aggregate(void) { return 1; }

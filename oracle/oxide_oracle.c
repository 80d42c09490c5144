/*
 * oxide_oracle.c — CPU restatement of OxideLake's CPU reference for the four
 * device operators (crates/oxidelake-device/src/cpu/kernels.rs).
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the MI355X kernels are
 * compared with. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it; the product library never links or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math, so
 * the float folds are the sequential IEEE folds the Rust reference computes:
 * Rust never contracts a*b+c into an FMA).
 *
 * Pinning (what fixes each function besides this restatement):
 *  - ora_total_order_key:  kernels.rs:1318-1358 round trip + ordering table
 *  - ora_eval_predicate, ora_eval_expr, ora_compact: kernels.rs:976-1021 known answers, the
 *    hand-written DataFusion order of conformance.rs:1460-1495, brute force
 *    on tiny inputs; extended grammar (Or, %, Int32/Float32, arithmetic):
 *    pinned by pyarrow compute cross-checks (integer and NaN-free cases) and
 *    closed forms; DataFusion's own coercion for it is parity unpinned.
 *  - ora_agg_*: kernels.rs:1123-1288, conformance.rs:306-356 overflow rows.
 *  - ora_join_*: kernels.rs:1033-1121, nested-loop brute force.
 *  - ora_l2/ora_cosine: kernels.rs:1360-1396 known answers, closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/oxide_rocm.h"

#define ORA_MAX_STACK 64

/* ------------------------------------------------------------ float order */

/* kernels.rs:825-830 `total_order_key`: f64 bits as an i64 that orders like
 * IEEE totalOrder (negative values get every bit but the sign flipped). */
int64_t ora_total_order_key(double x) {
  int64_t bits;
  memcpy(&bits, &x, 8);
  return bits ^ (int64_t)(((uint64_t)(bits >> 63)) >> 1);
}

/* kernels.rs:836-838 `from_total_order_key`: the transform is an involution. */
double ora_from_total_order_key(int64_t key) {
  int64_t bits = key ^ (int64_t)(((uint64_t)(key >> 63)) >> 1);
  double x;
  memcpy(&x, &bits, 8);
  return x;
}

static int32_t total_order_key32(float x) {
  int32_t bits;
  memcpy(&bits, &x, 4);
  return bits ^ (int32_t)(((uint32_t)(bits >> 31)) >> 1);
}

/* kernels.rs:62-69 `positive_zero`: -0.0 as +0.0, everything else as is. */
static double positive_zero(double v) { return v == 0.0 ? 0.0 : v; }
static float positive_zero32(float v) { return v == 0.0f ? 0.0f : v; }

/* ------------------------------------------------------------ expressions */

typedef struct {
  int32_t type;      /* OXR_INT32 .. OXR_FLOAT64, or 0 for boolean */
  uint8_t *valid;    /* n bytes 0/1 */
  int64_t *i;        /* ints (Int32 held sign-extended) / booleans */
  double *f;         /* Float64 */
  float *f32;        /* Float32 */
} ora_val;

static int is_float(int t) { return t == OXR_FLOAT32 || t == OXR_FLOAT64; }

/* DataFusion's numeric coercion for arithmetic and comparisons. */
int32_t ora_coerce(int32_t a, int32_t b) {
  if (a == OXR_FLOAT64 || b == OXR_FLOAT64) return OXR_FLOAT64;
  if (a == OXR_FLOAT32 || b == OXR_FLOAT32) return OXR_FLOAT32;
  if (a == OXR_INT64 || b == OXR_INT64) return OXR_INT64;
  return OXR_INT32;
}

static void val_alloc(ora_val *v, int32_t type, size_t n) {
  v->type = type;
  v->valid = (uint8_t *)calloc(n ? n : 1, 1);
  v->i = (int64_t *)calloc(n ? n : 1, 8);
  v->f = (double *)calloc(n ? n : 1, 8);
  v->f32 = (float *)calloc(n ? n : 1, 4);
}
static void val_free(ora_val *v) {
  free(v->valid);
  free(v->i);
  free(v->f);
  free(v->f32);
}

/* Casts row r of v to type t (values held in the slot for t). */
static void cast_row(const ora_val *v, size_t r, int32_t t, int64_t *oi, double *of, float *of32) {
  switch (v->type) {
  case OXR_INT32:
  case OXR_INT64:
    if (t == OXR_INT32) *oi = (int64_t)(int32_t)v->i[r];
    else if (t == OXR_INT64) *oi = v->i[r];
    else if (t == OXR_FLOAT64) *of = (double)v->i[r];
    else *of32 = (float)v->i[r];
    break;
  case OXR_FLOAT32:
    if (t == OXR_FLOAT64) *of = (double)v->f32[r];
    else *of32 = v->f32[r];
    break;
  default:
    *of = v->f[r];
    break;
  }
}

static int column_value(const oxr_column *c, size_t n, ora_val *out) {
  val_alloc(out, c->type, n);
  for (size_t r = 0; r < n; r++) {
    out->valid[r] = c->validity ? (c->validity[r >> 3] >> (r & 7)) & 1 : 1;
    switch (c->type) {
    case OXR_INT32: out->i[r] = ((const int32_t *)c->values)[r]; break;
    case OXR_INT64: out->i[r] = ((const int64_t *)c->values)[r]; break;
    case OXR_FLOAT32: out->f32[r] = ((const float *)c->values)[r]; break;
    case OXR_FLOAT64: out->f[r] = ((const double *)c->values)[r]; break;
    default: return -1;
    }
  }
  return 0;
}

/* a `op` b for one row, in the coerced type. Returns -1 for an ordering. */
static int compare_row(const ora_val *a, const ora_val *b, size_t r, int32_t t, int32_t op) {
  int c;
  if (t == OXR_FLOAT64) {
    double x = 0, y = 0;
    int64_t di;
    float df;
    cast_row(a, r, t, &di, &x, &df);
    cast_row(b, r, t, &di, &y, &df);
    int64_t kx = ora_total_order_key(positive_zero(x)), ky = ora_total_order_key(positive_zero(y));
    c = kx < ky ? -1 : kx > ky ? 1 : 0;
  } else if (t == OXR_FLOAT32) {
    float x = 0, y = 0;
    int64_t di;
    double dd;
    cast_row(a, r, t, &di, &dd, &x);
    cast_row(b, r, t, &di, &dd, &y);
    int32_t kx = total_order_key32(positive_zero32(x)), ky = total_order_key32(positive_zero32(y));
    c = kx < ky ? -1 : kx > ky ? 1 : 0;
  } else {
    int64_t x = 0, y = 0;
    double dd;
    float df;
    cast_row(a, r, t, &x, &dd, &df);
    cast_row(b, r, t, &y, &dd, &df);
    c = x < y ? -1 : x > y ? 1 : 0;
  }
  switch (op) {
  case OXR_EQ: return c == 0;
  case OXR_LT: return c < 0;
  case OXR_LTEQ: return c <= 0;
  case OXR_GT: return c > 0;
  case OXR_GTEQ: return c >= 0;
  case OXR_NOTEQ: return c != 0;
  }
  return 0;
}

/* Evaluates a postfix expression over n rows of ncols host columns. On
 * success *out holds the result (the caller frees it with ora_val_free).
 * Returns 0, or: 1 = division by zero in a valid row, 2 = integer overflow
 * in a division, -1 = malformed expression. */
static int eval_expr(const oxr_column *cols, int32_t ncols, size_t n, const oxr_expr *e,
                     ora_val *out) {
  ora_val st[ORA_MAX_STACK];
  int sp = 0, err = 0;
  for (int32_t k = 0; k < e->count; k++) {
    const oxr_expr_node *nd = &e->nodes[k];
    if (nd->kind == OXR_X_COLUMN) {
      if (sp >= ORA_MAX_STACK || nd->column < 0 || nd->column >= ncols) { err = -1; break; }
      if (column_value(&cols[nd->column], n, &st[sp])) { err = -1; break; }
      sp++;
    } else if (nd->kind == OXR_X_LITERAL) {
      if (sp >= ORA_MAX_STACK) { err = -1; break; }
      val_alloc(&st[sp], nd->type, n);
      for (size_t r = 0; r < n; r++) {
        st[sp].valid[r] = 1;
        if (nd->type == OXR_INT32) st[sp].i[r] = (int32_t)nd->lit_i;
        else if (nd->type == OXR_INT64) st[sp].i[r] = nd->lit_i;
        else if (nd->type == OXR_FLOAT32) st[sp].f32[r] = (float)nd->lit_f;
        else st[sp].f[r] = nd->lit_f;
      }
      sp++;
    } else {
      if (sp < 2) { err = -1; break; }
      ora_val *a = &st[sp - 2], *b = &st[sp - 1];
      ora_val res;
      if (nd->kind == OXR_X_AND || nd->kind == OXR_X_OR) {
        /* Kleene logic (arrow `and_kleene` / `or_kleene`). */
        val_alloc(&res, 0, n);
        for (size_t r = 0; r < n; r++) {
          int av = a->valid[r], bv = b->valid[r];
          int ax = (int)a->i[r], bx = (int)b->i[r];
          if (nd->kind == OXR_X_AND) {
            if ((av && !ax) || (bv && !bx)) { res.valid[r] = 1; res.i[r] = 0; }
            else if (av && bv) { res.valid[r] = 1; res.i[r] = 1; }
            else res.valid[r] = 0;
          } else {
            if ((av && ax) || (bv && bx)) { res.valid[r] = 1; res.i[r] = 1; }
            else if (av && bv) { res.valid[r] = 1; res.i[r] = 0; }
            else res.valid[r] = 0;
          }
        }
      } else if (nd->kind == OXR_X_CMP) {
        if (a->type == 0 || b->type == 0) { err = -1; break; }
        int32_t t = ora_coerce(a->type, b->type);
        val_alloc(&res, 0, n);
        for (size_t r = 0; r < n; r++) {
          res.valid[r] = a->valid[r] && b->valid[r];
          res.i[r] = res.valid[r] ? compare_row(a, b, r, t, nd->op) : 0;
        }
      } else {
        if (a->type == 0 || b->type == 0) { err = -1; break; }
        int32_t t = ora_coerce(a->type, b->type);
        val_alloc(&res, t, n);
        for (size_t r = 0; r < n && !err; r++) {
          res.valid[r] = a->valid[r] && b->valid[r];
          if (!res.valid[r]) continue;
          int64_t xi = 0, yi = 0;
          double xf = 0, yf = 0;
          float xs = 0, ys = 0;
          cast_row(a, r, t, &xi, &xf, &xs);
          cast_row(b, r, t, &yi, &yf, &ys);
          if (t == OXR_INT32 || t == OXR_INT64) {
            uint64_t ux = (uint64_t)xi, uy = (uint64_t)yi, z = 0;
            switch (nd->kind) {
            case OXR_X_ADD: z = ux + uy; break;
            case OXR_X_SUB: z = ux - uy; break;
            case OXR_X_MUL: z = ux * uy; break;
            case OXR_X_DIV:
            case OXR_X_MOD: {
              int64_t lo = t == OXR_INT32 ? INT32_MIN : INT64_MIN;
              if (yi == 0) { err = 1; break; }
              if (xi == lo && yi == -1) {
                if (nd->kind == OXR_X_DIV) { err = 2; break; }
                z = 0;
              } else {
                z = (uint64_t)(nd->kind == OXR_X_DIV ? xi / yi : xi % yi);
              }
              break;
            }
            default: err = -1;
            }
            res.i[r] = t == OXR_INT32 ? (int64_t)(int32_t)(uint32_t)z : (int64_t)z;
          } else if (t == OXR_FLOAT64) {
            double z = 0;
            switch (nd->kind) {
            case OXR_X_ADD: z = xf + yf; break;
            case OXR_X_SUB: z = xf - yf; break;
            case OXR_X_MUL: z = xf * yf; break;
            case OXR_X_DIV: z = xf / yf; break;
            case OXR_X_MOD: z = fmod(xf, yf); break;
            default: err = -1;
            }
            res.f[r] = z;
          } else {
            float z = 0;
            switch (nd->kind) {
            case OXR_X_ADD: z = xs + ys; break;
            case OXR_X_SUB: z = xs - ys; break;
            case OXR_X_MUL: z = xs * ys; break;
            case OXR_X_DIV: z = xs / ys; break;
            case OXR_X_MOD: z = fmodf(xs, ys); break;
            default: err = -1;
            }
            res.f32[r] = z;
          }
        }
      }
      val_free(a);
      val_free(b);
      sp -= 2;
      st[sp++] = res;
      if (err) break;
    }
  }
  if (!err && sp != 1) err = -1;
  if (err) {
    for (int i = 0; i < sp; i++) val_free(&st[i]);
    return err;
  }
  *out = st[0];
  return 0;
}

/* Result type of an expression without evaluating it (0 = boolean), or -1. */
int32_t ora_expr_type(const oxr_column *cols, int32_t ncols, const oxr_expr *e) {
  int32_t st[ORA_MAX_STACK];
  int sp = 0;
  for (int32_t k = 0; k < e->count; k++) {
    const oxr_expr_node *nd = &e->nodes[k];
    if (nd->kind == OXR_X_COLUMN) {
      if (sp >= ORA_MAX_STACK || nd->column < 0 || nd->column >= ncols) return -1;
      st[sp++] = cols[nd->column].type;
    } else if (nd->kind == OXR_X_LITERAL) {
      if (sp >= ORA_MAX_STACK) return -1;
      st[sp++] = nd->type;
    } else {
      if (sp < 2) return -1;
      int32_t a = st[sp - 2], b = st[sp - 1];
      sp -= 2;
      if (nd->kind == OXR_X_AND || nd->kind == OXR_X_OR) st[sp++] = 0;
      else if (nd->kind == OXR_X_CMP) st[sp++] = 0;
      else st[sp++] = ora_coerce(a, b);
    }
  }
  return sp == 1 ? st[0] : -1;
}

/* kernels.rs:99-141 `evaluate_predicate` (extended grammar): mask[r] = 1 when
 * the predicate is true (not NULL, not false) at row r. */
int ora_eval_predicate(const oxr_column *cols, int32_t ncols, uint64_t n, const oxr_expr *pred,
                       uint8_t *mask) {
  if (!pred || pred->count == 0) {
    memset(mask, 1, n);
    return 0;
  }
  ora_val v;
  int err = eval_expr(cols, ncols, n, pred, &v);
  if (err) return err;
  for (uint64_t r = 0; r < n; r++) mask[r] = v.valid[r] && v.i[r] != 0;
  val_free(&v);
  return 0;
}

/* One projection expression evaluated at every row: values written at the
 * type's width into `values`, validity as bytes. Returns its type via *type. */
int ora_eval_expr(const oxr_column *cols, int32_t ncols, uint64_t n, const oxr_expr *e,
                  void *values, uint8_t *valid, int32_t *type) {
  ora_val v;
  int err = eval_expr(cols, ncols, n, e, &v);
  if (err) return err;
  *type = v.type;
  for (uint64_t r = 0; r < n; r++) {
    valid[r] = v.valid[r];
    switch (v.type) {
    case OXR_INT32: ((int32_t *)values)[r] = (int32_t)v.i[r]; break;
    case OXR_INT64: ((int64_t *)values)[r] = v.i[r]; break;
    case OXR_FLOAT32: ((float *)values)[r] = v.f32[r]; break;
    case OXR_FLOAT64: ((double *)values)[r] = v.f[r]; break;
    default: ((int64_t *)values)[r] = v.i[r]; break;
    }
  }
  val_free(&v);
  return 0;
}

/* kernels.rs:144-152 `filter_project` (mask, then `filter_record_batch`)
 * for one output column: out[j] = value of `e` at the j-th selected row.
 * A single-column projection of a vector column copies dim floats per row.
 * Returns the number of selected rows, or a negative error. */
int64_t ora_compact(const uint8_t *mask, uint64_t n, const void *values, const uint8_t *valid,
                    size_t width, void *out, uint8_t *out_valid) {
  uint64_t j = 0;
  for (uint64_t r = 0; r < n; r++) {
    if (!mask[r]) continue;
    memcpy((char *)out + j * width, (const char *)values + r * width, width);
    if (out_valid) out_valid[j] = valid ? valid[r] : 1;
    j++;
  }
  return (int64_t)j;
}

/* ------------------------------------------------------------ aggregation */

/* kernels.rs:478-818 `GroupedAggregator`/`Accumulator`: streaming group
 * state. Groups open in first-seen order; NULL keys form one group; SUM,
 * MIN and MAX skip NULLs and stay NULL for a group without values; COUNT
 * counts non-NULL values (COUNT(*) counts rows); SUM over integers wraps in
 * Int64; SUM over floats is a Float64 fold from +0.0 (sum_mode 0), or a
 * Neumaier-compensated sum (sum_mode 1) used as the exact value in
 * tolerance checks; MIN/MAX of floats follow totalOrder. */
typedef struct {
  int32_t func;
  int32_t in_type;  /* input value type */
  int64_t *i;       /* per group: int64 sum/min/max, count */
  double *f;        /* per group: float sum, or Float64 min/max */
  double *comp;     /* Neumaier compensation */
  float *f32;       /* Float32 min/max */
  uint8_t *has;     /* per group: saw a value */
} ora_acc;

typedef struct {
  int32_t key_type;
  int32_t naggs;
  int32_t sum_mode;
  ora_acc *accs;
  /* open addressing key -> group */
  uint64_t cap;
  int64_t *slot_key;
  int64_t *slot_gid; /* -1 empty */
  int64_t null_gid;
  /* groups */
  int64_t ngroups, gcap;
  int64_t *keys;
  uint8_t *key_null;
} ora_agg;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

ora_agg *ora_agg_new(int32_t key_type, int32_t naggs, const int32_t *funcs, const int32_t *in_types,
                     int32_t sum_mode) {
  ora_agg *a = (ora_agg *)calloc(1, sizeof(ora_agg));
  a->key_type = key_type;
  a->naggs = naggs;
  a->sum_mode = sum_mode;
  a->accs = (ora_acc *)calloc(naggs ? naggs : 1, sizeof(ora_acc));
  for (int i = 0; i < naggs; i++) {
    a->accs[i].func = funcs[i];
    a->accs[i].in_type = in_types[i];
  }
  a->cap = 16;
  a->slot_key = (int64_t *)calloc(a->cap, 8);
  a->slot_gid = (int64_t *)malloc(a->cap * 8);
  for (uint64_t s = 0; s < a->cap; s++) a->slot_gid[s] = -1;
  a->null_gid = -1;
  return a;
}

static void agg_grow_groups(ora_agg *a) {
  if (a->ngroups < a->gcap) return;
  int64_t nc = a->gcap ? a->gcap * 2 : 16;
  a->keys = (int64_t *)realloc(a->keys, nc * 8);
  a->key_null = (uint8_t *)realloc(a->key_null, nc);
  for (int i = 0; i < a->naggs; i++) {
    ora_acc *c = &a->accs[i];
    c->i = (int64_t *)realloc(c->i, nc * 8);
    c->f = (double *)realloc(c->f, nc * 8);
    c->comp = (double *)realloc(c->comp, nc * 8);
    c->f32 = (float *)realloc(c->f32, nc * 4);
    c->has = (uint8_t *)realloc(c->has, nc);
  }
  a->gcap = nc;
}

static int64_t agg_open_group(ora_agg *a, int64_t key, int is_null) {
  agg_grow_groups(a);
  int64_t g = a->ngroups++;
  a->keys[g] = is_null ? 0 : key;
  a->key_null[g] = (uint8_t)is_null;
  for (int i = 0; i < a->naggs; i++) {
    ora_acc *c = &a->accs[i];
    c->i[g] = 0;
    c->f[g] = 0.0;
    c->comp[g] = 0.0;
    c->f32[g] = 0.0f;
    c->has[g] = 0;
  }
  return g;
}

static void agg_rehash(ora_agg *a) {
  uint64_t nc = a->cap * 2;
  int64_t *nk = (int64_t *)calloc(nc, 8), *ng = (int64_t *)malloc(nc * 8);
  for (uint64_t s = 0; s < nc; s++) ng[s] = -1;
  for (uint64_t s = 0; s < a->cap; s++) {
    if (a->slot_gid[s] < 0) continue;
    uint64_t h = mix64((uint64_t)a->slot_key[s]) & (nc - 1);
    while (ng[h] >= 0) h = (h + 1) & (nc - 1);
    nk[h] = a->slot_key[s];
    ng[h] = a->slot_gid[s];
  }
  free(a->slot_key);
  free(a->slot_gid);
  a->slot_key = nk;
  a->slot_gid = ng;
  a->cap = nc;
}

/* kernels.rs:738-762 `group_ids`. */
static int64_t agg_group_of(ora_agg *a, int64_t key, int is_null) {
  if (is_null) {
    if (a->null_gid < 0) a->null_gid = agg_open_group(a, 0, 1);
    return a->null_gid;
  }
  if ((uint64_t)(a->ngroups + 1) * 2 > a->cap) agg_rehash(a);
  uint64_t h = mix64((uint64_t)key) & (a->cap - 1);
  while (a->slot_gid[h] >= 0) {
    if (a->slot_key[h] == key) return a->slot_gid[h];
    h = (h + 1) & (a->cap - 1);
  }
  int64_t g = agg_open_group(a, key, 0);
  a->slot_key[h] = key;
  a->slot_gid[h] = g;
  return g;
}

static void acc_fold(ora_acc *c, int64_t g, int32_t sum_mode, int64_t vi, double vf, float vs,
                     int merge) {
  switch (c->func) {
  case OXR_COUNT:
  case OXR_COUNT_STAR:
    c->i[g] = (int64_t)((uint64_t)c->i[g] + (uint64_t)(merge ? vi : 1));
    c->has[g] = 1;
    return;
  case OXR_SUM:
    if (c->in_type == OXR_INT32 || c->in_type == OXR_INT64) {
      c->i[g] = c->has[g] ? (int64_t)((uint64_t)c->i[g] + (uint64_t)vi) : vi;
    } else {
      double v = c->in_type == OXR_FLOAT32 && !merge ? (double)vs : vf;
      if (!c->has[g]) {
        c->f[g] = 0.0 + v; /* kernels.rs:595-602: from 0.0, so SUM{-0.0} = 0.0 */
        c->comp[g] = 0.0;
      } else if (sum_mode == 0) {
        c->f[g] = c->f[g] + v;
      } else {
        double s = c->f[g], t = s + v;
        if (isfinite(t)) {
          if (fabs(s) >= fabs(v)) c->comp[g] += (s - t) + v;
          else c->comp[g] += (v - t) + s;
        }
        c->f[g] = t;
      }
    }
    c->has[g] = 1;
    return;
  case OXR_MIN:
  case OXR_MAX: {
    int mx = c->func == OXR_MAX;
    if (c->in_type == OXR_INT32 || c->in_type == OXR_INT64) {
      if (!c->has[g] || (mx ? vi > c->i[g] : vi < c->i[g])) c->i[g] = vi;
    } else if (c->in_type == OXR_FLOAT64) {
      /* kernels.rs:527-538 total_min / total_max */
      if (!c->has[g]) c->f[g] = vf;
      else {
        int64_t kv = ora_total_order_key(vf), kh = ora_total_order_key(c->f[g]);
        if (mx ? kv > kh : kv < kh) c->f[g] = vf;
      }
    } else {
      if (!c->has[g]) c->f32[g] = vs;
      else {
        int32_t kv = total_order_key32(vs), kh = total_order_key32(c->f32[g]);
        if (mx ? kv > kh : kv < kh) c->f32[g] = vs;
      }
    }
    c->has[g] = 1;
    return;
  }
  }
}

/* kernels.rs:765-772 `update`: folds n input rows. keys/key_valid are the
 * group key (Int64 or Int32 widened); vals[i] / val_valid[i] the i-th
 * aggregate's input (int64 for integer inputs, double for Float64, float for
 * Float32; ignored for COUNT(*)). Validity arrays are bytes (NULL = all). */
void ora_agg_update(ora_agg *a, uint64_t n, const int64_t *keys, const uint8_t *key_valid,
                    const void *const *vals, const uint8_t *const *val_valid) {
  for (uint64_t r = 0; r < n; r++) {
    int knull = key_valid && !key_valid[r];
    int64_t g = agg_group_of(a, knull ? 0 : keys[r], knull);
    for (int i = 0; i < a->naggs; i++) {
      ora_acc *c = &a->accs[i];
      if (c->func == OXR_COUNT_STAR) {
        acc_fold(c, g, a->sum_mode, 0, 0, 0, 0);
        continue;
      }
      if (val_valid && val_valid[i] && !val_valid[i][r]) {
        if (c->func == OXR_COUNT) c->has[g] = 1;
        continue;
      }
      int64_t vi = 0;
      double vf = 0;
      float vs = 0;
      if (c->in_type == OXR_INT32 || c->in_type == OXR_INT64) vi = ((const int64_t *)vals[i])[r];
      else if (c->in_type == OXR_FLOAT64) vf = ((const double *)vals[i])[r];
      else vs = ((const float *)vals[i])[r];
      acc_fold(c, g, a->sum_mode, vi, vf, vs, 0);
    }
  }
}

/* kernels.rs:776-790 `merge`: folds a partial laid out as this aggregation's
 * output. Partial COUNTs add; SUM/MIN/MAX fold like input (a SUM over floats
 * arrives as Float64). */
void ora_agg_merge(ora_agg *a, uint64_t n, const int64_t *keys, const uint8_t *key_valid,
                   const void *const *vals, const uint8_t *const *val_valid) {
  for (uint64_t r = 0; r < n; r++) {
    int knull = key_valid && !key_valid[r];
    int64_t g = agg_group_of(a, knull ? 0 : keys[r], knull);
    for (int i = 0; i < a->naggs; i++) {
      ora_acc *c = &a->accs[i];
      if (val_valid && val_valid[i] && !val_valid[i][r]) continue;
      int64_t vi = 0;
      double vf = 0;
      float vs = 0;
      int int_in = c->in_type == OXR_INT32 || c->in_type == OXR_INT64;
      if (c->func == OXR_COUNT || c->func == OXR_COUNT_STAR || int_in)
        vi = ((const int64_t *)vals[i])[r];
      else if (c->func == OXR_SUM || c->in_type == OXR_FLOAT64) vf = ((const double *)vals[i])[r];
      else vs = ((const float *)vals[i])[r];
      acc_fold(c, g, a->sum_mode, vi, vf, vs, 1);
    }
  }
}

int64_t ora_agg_groups(const ora_agg *a) { return a->ngroups; }

static const ora_agg *g_sort_agg;
static int cmp_groups(const void *x, const void *y) {
  int64_t gx = *(const int64_t *)x, gy = *(const int64_t *)y;
  const ora_agg *a = g_sort_agg;
  if (a->key_null[gx] != a->key_null[gy]) return a->key_null[gx] ? 1 : -1;
  if (a->keys[gx] < a->keys[gy]) return -1;
  if (a->keys[gx] > a->keys[gy]) return 1;
  return 0;
}

/* kernels.rs:793-811 `finish`: groups sorted by key, NULL last. Writes
 * ngroups rows: keys/key_valid, and per aggregate i out_i[i] (int64 for
 * COUNT and integer SUM/MIN/MAX, double for float SUM and Float64 MIN/MAX,
 * float for Float32 MIN/MAX) with out_valid[i]. */
void ora_agg_finish(ora_agg *a, int64_t *keys, uint8_t *key_valid, void *const *outs,
                    uint8_t *const *out_valid) {
  int64_t *order = (int64_t *)malloc((a->ngroups ? a->ngroups : 1) * 8);
  for (int64_t g = 0; g < a->ngroups; g++) order[g] = g;
  g_sort_agg = a;
  qsort(order, a->ngroups, 8, cmp_groups);
  for (int64_t j = 0; j < a->ngroups; j++) {
    int64_t g = order[j];
    keys[j] = a->keys[g];
    key_valid[j] = !a->key_null[g];
    for (int i = 0; i < a->naggs; i++) {
      ora_acc *c = &a->accs[i];
      int int_in = c->in_type == OXR_INT32 || c->in_type == OXR_INT64;
      int is_count = c->func == OXR_COUNT || c->func == OXR_COUNT_STAR;
      out_valid[i][j] = is_count ? 1 : c->has[g];
      if (is_count || int_in) ((int64_t *)outs[i])[j] = c->i[g];
      else if (c->func == OXR_SUM) ((double *)outs[i])[j] = c->f[g] + c->comp[g];
      else if (c->in_type == OXR_FLOAT64) ((double *)outs[i])[j] = c->f[g];
      else ((float *)outs[i])[j] = c->f32[g];
    }
  }
  free(order);
}

void ora_agg_free(ora_agg *a) {
  if (!a) return;
  for (int i = 0; i < a->naggs; i++) {
    free(a->accs[i].i);
    free(a->accs[i].f);
    free(a->accs[i].comp);
    free(a->accs[i].f32);
    free(a->accs[i].has);
  }
  free(a->accs);
  free(a->slot_key);
  free(a->slot_gid);
  free(a->keys);
  free(a->key_null);
  free(a);
}

/* ------------------------------------------------------------------ join */

/* kernels.rs:199-226 `JoinTable::build`: every distinct non-NULL key's run
 * in a row array grouped by key, build order within a key (pass 1 counts,
 * pass 2 offsets, pass 3 scatters). Run placement order is this oracle's
 * own (the reference's follows HashMap iteration order); only the runs'
 * contents are observable. */
typedef struct {
  uint64_t cap;
  int64_t *slot_key;
  int64_t *slot_start; /* -1 = empty */
  int64_t *slot_count;
  uint32_t *rows;
  uint64_t nrows, distinct;
} ora_join;

static uint64_t join_find(const ora_join *t, int64_t key) {
  uint64_t h = mix64((uint64_t)key) & (t->cap - 1);
  while (t->slot_start[h] >= 0) {
    if (t->slot_key[h] == key) return h;
    h = (h + 1) & (t->cap - 1);
  }
  return h;
}

ora_join *ora_join_build(uint64_t n, const int64_t *keys, const uint8_t *valid) {
  ora_join *t = (ora_join *)calloc(1, sizeof(ora_join));
  t->cap = 16;
  while (t->cap < 2 * n) t->cap <<= 1;
  t->slot_key = (int64_t *)calloc(t->cap, 8);
  t->slot_start = (int64_t *)malloc(t->cap * 8);
  t->slot_count = (int64_t *)calloc(t->cap, 8);
  for (uint64_t s = 0; s < t->cap; s++) t->slot_start[s] = -1;
  /* pass 1: rows per key (runs placed in first-appearance order) */
  uint64_t next = 0;
  for (uint64_t j = 0; j < n; j++) {
    if (valid && !valid[j]) continue;
    uint64_t h = join_find(t, keys[j]);
    if (t->slot_start[h] < 0) {
      t->slot_key[h] = keys[j];
      t->slot_start[h] = 0;
      t->distinct++;
    }
    t->slot_count[h]++;
  }
  /* pass 2: offsets, runs laid out in order of each key's first row */
  int64_t *cursor = (int64_t *)calloc(t->cap, 8);
  for (uint64_t j = 0; j < n; j++) {
    if (valid && !valid[j]) continue;
    uint64_t h = join_find(t, keys[j]);
    if (cursor[h] == 0) {
      t->slot_start[h] = (int64_t)next;
      next += (uint64_t)t->slot_count[h];
      cursor[h] = 1;
    }
  }
  /* pass 3: scatter rows in build order */
  t->nrows = next;
  t->rows = (uint32_t *)malloc((next ? next : 1) * 4);
  memset(cursor, 0, t->cap * 8);
  for (uint64_t j = 0; j < n; j++) {
    if (valid && !valid[j]) continue;
    uint64_t h = join_find(t, keys[j]);
    t->rows[t->slot_start[h] + cursor[h]++] = (uint32_t)j;
  }
  free(cursor);
  return t;
}

uint64_t ora_join_distinct(const ora_join *t) { return t->distinct; }
uint64_t ora_join_rows(const ora_join *t) { return t->nrows; }

/* kernels.rs:320-342 `JoinMatches::probe`: offsets[i] = first match of probe
 * row i (n+1 entries, u64), starts[i] = its run's start. Returns the total. */
uint64_t ora_join_probe(const ora_join *t, uint64_t n, const int64_t *keys, const uint8_t *valid,
                        uint64_t *offsets, uint32_t *starts) {
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; i++) {
    offsets[i] = total;
    starts[i] = 0;
    if (valid && !valid[i]) continue;
    uint64_t h = join_find(t, keys[i]);
    if (t->slot_start[h] < 0) continue;
    starts[i] = (uint32_t)t->slot_start[h];
    total += (uint64_t)t->slot_count[h];
  }
  offsets[n] = total;
  return total;
}

/* kernels.rs:364-405 `JoinMatches::pairs`: (probe row, build row) of
 * matches [first, first+rows). Returns 0, or -1 when out of range. */
int ora_join_pairs(const ora_join *t, uint64_t n, const uint64_t *offsets, const uint32_t *starts,
                   uint64_t first, uint64_t rows, uint32_t *left, uint32_t *right) {
  uint64_t end = first + rows;
  if (end < first || end > offsets[n]) return -1;
  /* the probe row holding `first`: the last whose offset is <= first */
  uint64_t lo = 0, hi = n + 1;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (offsets[mid] <= first) lo = mid + 1;
    else hi = mid;
  }
  uint64_t row = lo ? lo - 1 : 0, at = first, k = 0;
  while (at < end) {
    uint64_t a = offsets[row], b = offsets[row + 1];
    uint64_t upto = b < end ? b : end;
    for (; at < upto; at++, k++) {
      left[k] = (uint32_t)row;
      right[k] = t->rows[starts[row] + (at - a)];
    }
    row++;
  }
  return 0;
}

void ora_join_free(ora_join *t) {
  if (!t) return;
  free(t->slot_key);
  free(t->slot_start);
  free(t->slot_count);
  free(t->rows);
  free(t);
}

/* --------------------------------------------------------- vector distance */

/* kernels.rs:908-915 `l2`: sqrt of an f32 fold of (x-y)^2. Rust's
 * `Sum for f32` folds from -0.0. */
float ora_l2(const float *a, const float *b, int32_t dim) {
  float s = -0.0f;
  for (int32_t i = 0; i < dim; i++) {
    float d = a[i] - b[i];
    s = s + d * d;
  }
  return sqrtf(s);
}

/* kernels.rs:918-927 `cosine`: 1 - dot/(|a||b|), NaN when a norm is 0. */
float ora_cosine(const float *a, const float *b, int32_t dim) {
  float dot = -0.0f, na = -0.0f, nb = -0.0f;
  for (int32_t i = 0; i < dim; i++) dot = dot + a[i] * b[i];
  for (int32_t i = 0; i < dim; i++) na = na + a[i] * a[i];
  for (int32_t i = 0; i < dim; i++) nb = nb + b[i] * b[i];
  na = sqrtf(na);
  nb = sqrtf(nb);
  if (na == 0.0f || nb == 0.0f) return NAN;
  return 1.0f - dot / (na * nb);
}

/* kernels.rs:841-905 `vector_distance`: one distance per row (n x dim row
 * major); rows whose valid byte is 0 get valid_out 0. */
void ora_vector_distance(const float *vecs, const uint8_t *valid, uint64_t n, int32_t dim,
                         const float *query, int32_t metric, float *out, uint8_t *valid_out) {
  for (uint64_t r = 0; r < n; r++) {
    int ok = valid ? valid[r] : 1;
    valid_out[r] = (uint8_t)ok;
    if (!ok) {
      out[r] = 0.0f;
      continue;
    }
    const float *v = vecs + r * (uint64_t)dim;
    out[r] = metric == OXR_COSINE ? ora_cosine(v, query, dim) : ora_l2(v, query, dim);
  }
}

/* FP64 distance of one row (the precision oracle): L2 or cosine. */
static double dist64(const float *v, const float *q, int32_t dim, int32_t metric) {
  if (metric == OXR_COSINE) {
    double dot = 0, na = 0, nb = 0;
    for (int32_t i = 0; i < dim; i++) {
      dot += (double)v[i] * q[i];
      na += (double)v[i] * v[i];
      nb += (double)q[i] * q[i];
    }
    if (na == 0 || nb == 0) return NAN;
    return 1.0 - dot / (sqrt(na) * sqrt(nb));
  }
  double s = 0;
  for (int32_t i = 0; i < dim; i++) {
    double d = (double)v[i] - q[i];
    s += d * d;
  }
  return sqrt(s);
}

/* Exact top-k by FP64 distance, ties by row ascending; NULL and NaN rows
 * are skipped. out_rows[k], out_dist[k]; missing entries get row -1. */
void ora_topk64(const float *vecs, const uint8_t *valid, uint64_t n, int32_t dim,
                const float *query, int32_t metric, int32_t k, int64_t *out_rows,
                double *out_dist) {
  int32_t have = 0;
  for (int32_t j = 0; j < k; j++) {
    out_rows[j] = -1;
    out_dist[j] = INFINITY;
  }
  for (uint64_t r = 0; r < n; r++) {
    if (valid && !valid[r]) continue;
    double d = dist64(vecs + r * (uint64_t)dim, query, dim, metric);
    if (isnan(d)) continue;
    if (have == k && !(d < out_dist[k - 1])) continue;
    int32_t j = have < k ? have++ : k - 1;
    while (j > 0 && out_dist[j - 1] > d) {
      out_dist[j] = out_dist[j - 1];
      out_rows[j] = out_rows[j - 1];
      j--;
    }
    out_dist[j] = d;
    out_rows[j] = (int64_t)r;
  }
}

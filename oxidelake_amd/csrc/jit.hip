// Generation, compilation (hipRTC) and caching of per-query filter kernels.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "jit.h"
// kJitHeaderCount, kJitHeaderNames, kJitHeaderText: the device headers the
// generated source includes (divlit.h, scalar.h, compact.h, filter_jit.h),
// embedded at build time.
#include "jit_sources.inc"

namespace oxr {
namespace {

bool is_int_type(int32_t t) { return t == OXR_INT32 || t == OXR_INT64; }

const char *cname(int32_t t) {
  switch (t) {
    case OXR_INT32: return "int32_t";
    case OXR_INT64: return "int64_t";
    case OXR_FLOAT32: return "float";
    default: return "double";
  }
}

const char *tprefix(int32_t t) {
  switch (t) {
    case OXR_INT32: return "i32";
    case OXR_INT64: return "i64";
    case OXR_FLOAT32: return "f32";
    default: return "f64";
  }
}

const char *cmp_op(int op) {
  switch (op) {
    case OXR_EQ: return "==";
    case OXR_LT: return "<";
    case OXR_LTEQ: return "<=";
    case OXR_GT: return ">";
    case OXR_GTEQ: return ">=";
    default: return "!=";
  }
}

// A generated value: C++ expression (a temporary or a row field), its SQL
// type, and a C++ boolean expression that is true where it is not NULL.
struct Val {
  std::string v;
  int32_t t;
  std::string ok;
  bool nullable;
};

std::string conj(const std::string &a, const std::string &b) {
  if (a == "true") return b;
  if (b == "true") return a;
  if (a == b) return a;
  return "(" + a + " && " + b + ")";
}

// Emits straight-line C++ for value and boolean subtrees of one expression.
// The typing and coercion rules are those of the interpreter's compiler
// (expr.hip): DataFusion's numeric coercion and widening casts only.
struct Gen {
  const std::vector<DevCol> &cols;
  JitFilter *jf;
  std::vector<char> &val_used;    // column values the row struct holds
  std::vector<char> &valid_used;  // column validity words it holds
  const oxr_expr *e = nullptr;
  std::string code;
  int tmp = 0;
  bool unsupported = false;
  // prepared mode: literals are kernel arguments L[i] (their bits), so one
  // binary serves every value; else they are constants of the code
  std::vector<uint64_t> *params = nullptr;

  std::string fresh() { return "t" + std::to_string(tmp++); }
  void stmt(const std::string &s) { code += "    " + s + "\n"; }

  int32_t type_of(int begin, int end) {
    const oxr_expr_node &nd = e->nodes[end - 1];
    if (nd.kind == OXR_X_COLUMN) return cols[nd.column].type;
    if (nd.kind == OXR_X_LITERAL) return nd.type;
    const int mid = split_operands(e, begin, end);
    return coerce(type_of(begin, mid), type_of(mid, end - 1));
  }

  std::string name_ok(const std::string &ok) {
    if (ok == "true" || ok == "in" || ok.find("&&") == std::string::npos) return ok;
    const std::string n = fresh();
    stmt("const bool " + n + " = " + ok + ";");
    return n;
  }

  Val cast(const Val &a, int32_t to) {
    if (a.t == to) return a;
    Val r = a;
    r.t = to;
    if (!(to == OXR_INT64 && a.t == OXR_INT32) && !(to == OXR_FLOAT32 && is_int_type(a.t)) &&
        !(to == OXR_FLOAT64 && (is_int_type(a.t) || a.t == OXR_FLOAT32))) {
      unsupported = true;  // never narrows; the interpreter reports it
      return r;
    }
    const std::string n = fresh();
    stmt(std::string("const ") + cname(to) + " " + n + " = (" + cname(to) + ")" + a.v + ";");
    r.v = n;
    return r;
  }

  // Literals are compile-time constants of the generated code (their bit
  // patterns, so every value, NaN payloads included, is exact): the
  // compiler folds them into the comparisons and arithmetic.
  Val literal(const oxr_expr_node &nd) {
    char buf[64];
    if (params) {
      if ((int)params->size() >= kJitLits) {
        unsupported = true;
        return {"0", nd.type, "true", false};
      }
      uint64_t bits = 0;
      const std::string ref = "L[" + std::to_string(params->size()) + "]";
      std::string v;
      if (nd.type == OXR_INT32) {
        bits = (uint64_t)(int64_t)(int32_t)nd.lit_i;
        v = "((int32_t)(int64_t)" + ref + ")";
      } else if (nd.type == OXR_INT64) {
        bits = (uint64_t)nd.lit_i;
        v = "((int64_t)" + ref + ")";
      } else if (nd.type == OXR_FLOAT32) {
        const float f = (float)nd.lit_f;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        bits = u;
        v = "as_f32(" + ref + ")";
      } else {
        std::memcpy(&bits, &nd.lit_f, 8);
        v = "as_f64(" + ref + ")";
      }
      params->push_back(bits);
      return {v, nd.type, "true", false};
    }
    if (nd.type == OXR_INT32) {
      std::snprintf(buf, sizeof buf, "((int32_t)0x%08xu)", (uint32_t)(int32_t)nd.lit_i);
    } else if (nd.type == OXR_INT64) {
      std::snprintf(buf, sizeof buf, "((int64_t)0x%016llxull)", (unsigned long long)nd.lit_i);
    } else if (nd.type == OXR_FLOAT32) {
      const float f = (float)nd.lit_f;
      uint32_t u;
      std::memcpy(&u, &f, 4);
      std::snprintf(buf, sizeof buf, "as_f32(0x%08xull)", u);
    } else {
      uint64_t u;
      std::memcpy(&u, &nd.lit_f, 8);
      std::snprintf(buf, sizeof buf, "as_f64(0x%016llxull)", (unsigned long long)u);
    }
    return {buf, nd.type, "true", false};
  }

  Val value(int begin, int end) {
    const oxr_expr_node &nd = e->nodes[end - 1];
    if (nd.kind == OXR_X_COLUMN) {
      const int c = nd.column;
      const DevCol &col = cols[c];
      val_used[c] = 1;
      const bool nb = col.validity != nullptr;
      if (nb) valid_used[c] = 1;
      const std::string cs = std::to_string(c);
      return {"x.c" + cs, col.type, nb ? "(in && row_bit(x.w" + cs + ", j))" : "in", nb};
    }
    if (nd.kind == OXR_X_LITERAL) return literal(nd);
    const int mid = split_operands(e, begin, end);
    const oxr_expr_node &rhs = e->nodes[mid];
    const bool divmod = nd.kind == OXR_X_DIV || nd.kind == OXR_X_MOD;
    const char *mod = nd.kind == OXR_X_MOD ? "true" : "false";
    const int64_t dlit = rhs.type == OXR_INT32 ? (int64_t)(int32_t)rhs.lit_i : rhs.lit_i;
    if (divmod && end - 1 - mid == 1 && rhs.kind == OXR_X_LITERAL && is_int_type(rhs.type) &&
        dlit != 0 && dlit != -1 && is_int_type(type_of(begin, mid)) && !params) {
      // integer division by a literal other than 0 and -1 can neither fail
      // nor overflow: C's truncating / and % are the SQL result, and the
      // compiler turns the constant divisor into a multiply-high
      const Val a = value(begin, mid);
      const int32_t t = coerce(a.t, rhs.type);
      const Val ca = cast(a, t);
      const Val d = cast(literal(rhs), t);
      const std::string n = fresh();
      stmt(std::string("const ") + cname(t) + " " + n + " = (" + cname(t) + ")(" + ca.v +
           (nd.kind == OXR_X_DIV ? " / " : " % ") + d.v + ");");
      return {n, t, ca.ok, ca.nullable};
    }
    const Val a = value(begin, mid);
    const Val b = value(mid, end - 1);
    const int32_t t = coerce(a.t, b.t);
    const Val ca = cast(a, t), cb = cast(b, t);
    const std::string ok = name_ok(conj(ca.ok, cb.ok));
    std::string expr;
    const std::string args = ca.v + ", " + cb.v;
    switch (nd.kind) {
      case OXR_X_ADD: expr = std::string(tprefix(t)) + "_add(" + args + ")"; break;
      case OXR_X_SUB: expr = std::string(tprefix(t)) + "_sub(" + args + ")"; break;
      case OXR_X_MUL: expr = std::string(tprefix(t)) + "_mul(" + args + ")"; break;
      case OXR_X_DIV:
      case OXR_X_MOD:
        if (is_int_type(t))
          expr = std::string(tprefix(t)) + "_divmod(" + mod + ", " + args + ", " + ok + ", &err)";
        else
          expr = std::string(tprefix(t)) + (nd.kind == OXR_X_DIV ? "_div(" : "_mod(") + args + ")";
        break;
      default:
        unsupported = true;
        return {"0", t, "true", false};
    }
    const std::string n = fresh();
    stmt(std::string("const ") + cname(t) + " " + n + " = " + expr + ";");
    return {n, t, ok, ca.nullable || cb.nullable};
  }

  // Kleene AND/OR over comparisons with NULL read as false keep exactly the
  // rows the three-valued result keeps (there is no NOT in the grammar).
  std::string boolean(int begin, int end) {
    const oxr_expr_node &nd = e->nodes[end - 1];
    const int mid = split_operands(e, begin, end);
    const std::string n = fresh();
    if (nd.kind == OXR_X_CMP) {
      const Val a = value(begin, mid);
      const Val b = value(mid, end - 1);
      const int32_t t = coerce(a.t, b.t);
      const Val ca = cast(a, t), cb = cast(b, t);
      std::string cmp;
      if (t == OXR_FLOAT64)
        cmp = "order_key64(" + ca.v + ") " + cmp_op(nd.op) + " order_key64(" + cb.v + ")";
      else if (t == OXR_FLOAT32)
        cmp = "order_key32(" + ca.v + ") " + cmp_op(nd.op) + " order_key32(" + cb.v + ")";
      else
        cmp = ca.v + " " + cmp_op(nd.op) + " " + cb.v;
      const std::string ok = conj(ca.ok, cb.ok);
      stmt("const bool " + n + " = " + (ok == "true" ? cmp : ok + " && (" + cmp + ")") + ";");
      return n;
    }
    const std::string l = boolean(begin, mid);
    const std::string r = boolean(mid, end - 1);
    stmt("const bool " + n + " = " + l + (nd.kind == OXR_X_AND ? " & " : " | ") + r + ";");
    return n;
  }
};

// The input struct (column pointers advanced to the tile's first row), the
// row struct and its loads, for the columns the generated code reads.
void row_code(const std::vector<DevCol> &cols, const std::vector<char> &val_used,
              const std::vector<char> &valid_used, std::string *in_fields, std::string *in_init,
              std::string *row, std::string *load, int *vgprs) {
  *row = "  struct Row {\n";
  // column values are read once: nontemporal loads (knob jit.stream_loads,
  // default on) stream them past the caches (16-byte reads measured 7.0 TB/s
  // so against 6.3 with plain loads)
  const std::string ld = tune_get("jit.stream_loads", 1) > 0 ? "ld_stream" : "ld_plain";
  for (size_t c = 0; c < cols.size(); c++) {
    const std::string cs = std::to_string(c);
    if (val_used[c]) {
      const int32_t t = cols[c].type;
      *in_fields += std::string("    const ") + cname(t) + " *c" + cs + ";\n";
      *in_init += std::string("    i.c") + cs + " = (const " + cname(t) + " *)p.col[" + cs + "] + base;\n";
      *row += std::string("    ") + cname(t) + " c" + cs + ";\n";
      *load += std::string("    x.c") + cs + " = in ? " + ld + "(&i.c" + cs + "[j]) : (" + cname(t) + ")0;\n";
      *vgprs += (t == OXR_INT64 || t == OXR_FLOAT64) ? 2 : 1;
    }
    if (valid_used[c]) {
      *in_fields += "    const uint32_t *w" + cs + ";\n";
      *in_init += "    i.w" + cs + " = p.valid[" + cs + "] + (base >> 5);\n";
      *row += "    uint32_t w" + cs + ";\n";
      *load += "    x.w" + cs + " = in ? " + ld + "(&i.w" + cs + "[j >> 5]) : 0u;\n";
      *vgprs += 1;
    }
  }
  *row += "  };\n";
}

// S (rows per lane) for a row of `vgprs` 32-bit registers: the largest power
// of two up to 16 keeping the tile's registers near 128, unless the
// "filter.rows" knob says otherwise.
// Rows per lane of a generated kernel: the knob when set (oxr_tune_set only
// accepts 1-32, what the skeletons take), else the largest power of two up to
// 16 whose rows fit about 128 VGPRs.
int pick_rows(int vgprs, const char *knob_name = "filter.rows") {
  const int64_t knob = tune_get(knob_name, 0);
  if (knob >= 1 && knob <= kMaxRowsPerLane) return (int)knob;
  int s = 16;
  while (s > 2 && s * vgprs > 128) s /= 2;
  return s;
}

int generate(const std::vector<DevCol> &cols, const oxr_expr *pred, const oxr_expr *proj,
             int32_t nproj, const std::vector<ValueProg> &vps, JitFilter *jf) {
  if (cols.size() > (size_t)kJitCols || nproj > kJitOuts) return OXR_EUNSUPPORTED;
  std::vector<char> val_used(cols.size(), 0), valid_used(cols.size(), 0);
  Gen g{cols, jf, val_used, valid_used};
  if (tune_get("jit.params", 0) > 0) g.params = &jf->lits;
  // predicate
  std::string pred_code = "    return true;\n";
  if (pred && pred->count > 0) {
    g.e = pred;
    g.code.clear();
    g.tmp = 0;
    const std::string res = g.boolean(0, pred->count);
    pred_code = g.code + "    return " + res + ";\n";
  }
  // the staged kernel (knob filter.staged, emitted beside the tile kernel)
  // parks kept rows' projections in LDS; vector projections keep the tile
  // kernel's whole-wave copies
  bool has_vec = false;
  for (int k = 0; k < nproj; k++)
    if (proj[k].count == 1 && proj[k].nodes[0].kind == OXR_X_COLUMN && proj[k].nodes[0].column >= 0 &&
        proj[k].nodes[0].column < (int)cols.size() && cols[proj[k].nodes[0].column].type == OXR_F32_VECTOR)
      has_vec = true;
  const bool persistent = tune_get("filter.persistent", 0) > 0;
  const bool staged = tune_get("filter.staged", 1) > 0 && !persistent && !has_vec;
  const bool stream_stores = tune_get("jit.stream_stores", 0) > 0;
  // the staged kernel's copy out of LDS writes whole lines: nontemporal
  // (knob filter.stage_nt; 1.19 -> 1.11-1.13 ms on the 200M-row query)
  const bool stage_nt = tune_get("filter.stage_nt", 1) > 0;
  // projections
  std::string proj_code = "    const bool in = true;\n    (void)in;\n";
  std::string vec_code, out_fields, out_init, stage_init, copy_code;
  int nvalid = 0, nvec = 0, stage_bytes = 0;
  jf->vslot.assign(nproj, -1);
  for (int k = 0; k < nproj; k++) {
    const oxr_expr &pe = proj[k];
    const std::string ks = std::to_string(k);
    if (vps[k].nullable) jf->vslot[k] = nvalid++;
    const std::string setbit =
        "set_out_bit(bits + " + std::to_string(jf->vslot[k]) + " * stride, bitpos);";
    if (pe.count == 1 && pe.nodes[0].kind == OXR_X_COLUMN && cols[pe.nodes[0].column].type == OXR_F32_VECTOR) {
      const int c = pe.nodes[0].column;
      const std::string cs = std::to_string(c);
      out_fields += "    const float *v" + ks + ";\n    float *o" + ks + ";\n    int d" + ks + ";\n";
      out_init += "    o.d" + ks + " = p.dim[" + ks + "];\n    o.v" + ks + " = (const float *)p.col[" + cs +
                  "] + base * (uint64_t)o.d" + ks + ";\n    o.o" + ks + " = (float *)p.out[" + ks +
                  "] + at * (uint64_t)o.d" + ks + ";\n";
      vec_code += "    copy_vector_rows(o.v" + ks + ", o.o" + ks + ", o.d" + ks + ", sel, j, q, lane);\n";
      nvec++;
      if (vps[k].nullable) {
        valid_used[c] = 1;
        proj_code += "    if (row_bit(x.w" + cs + ", j)) " + setbit + "\n";
      }
      continue;
    }
    g.e = &pe;
    g.code.clear();
    const Val v = g.value(0, pe.count);
    if (v.t != vps[k].type) return OXR_EUNSUPPORTED;
    proj_code += "    {\n" + g.code;
    out_fields += std::string("    ") + cname(v.t) + " *o" + ks + ";\n";
    out_init += std::string("    o.o") + ks + " = (" + cname(v.t) + " *)p.out[" + ks + "] + at;\n";
    // staged: the stage's array for output k starts at cap * (bytes of the
    // outputs before it)
    const std::string tk = cname(v.t);
    stage_init += "    o.o" + ks + " = (" + tk + " *)(stage + (uint64_t)cap * " + std::to_string(stage_bytes) + ");\n";
    copy_code += "    {\n      const " + tk + " *t = (const " + tk + " *)(stage + (uint64_t)cap * " +
                 std::to_string(stage_bytes) + ");\n      for (uint32_t i = (uint32_t)tid; i < n; i += (uint32_t)nt) " +
                 (stage_nt ? "st_stream(&o.o" + ks + "[i], t[i]);" : "o.o" + ks + "[i] = t[i];") + "\n    }\n";
    stage_bytes += (v.t == OXR_INT32 || v.t == OXR_FLOAT32) ? 4 : 8;
    if (stream_stores)
      proj_code += "    st_stream(&o.o" + ks + "[q], " + v.v + ");\n";
    else
      proj_code += "    o.o" + ks + "[q] = " + v.v + ";\n";
    if (vps[k].nullable) proj_code += "    if (" + v.ok + ") " + setbit + "\n";
    proj_code += "    }\n";
  }
  if (g.unsupported) return OXR_EUNSUPPORTED;
  std::string in_fields, in_init, row, load;
  int vgprs = 0;
  row_code(cols, val_used, valid_used, &in_fields, &in_init, &row, &load, &vgprs);
  // the persistent kernel (knob filter.persistent, default off: 3x slower
  // on the 200M-row query, its look-backs walk back through whole rounds of
  // tiles compacted in step) holds two tiles' rows: half the rows per lane
  jf->persistent = persistent;
  // waves per workgroup (knob filter.waves): a tile of 64 * W * S rows, so
  // more waves mean fewer tiles and fewer look-backs for the same registers
  jf->waves = (int)tune_get("filter.waves", 4);
  jf->rows_per_lane = pick_rows(std::max(vgprs, 1) * (jf->persistent ? 2 : 1));
  // staged: W' waves, chunks of C rows per lane, NCH chunks per tile, a
  // stage of CAP kept rows (knobs filter.staged_waves, filter.chunk_rows,
  // filter.chunks, filter.stage_rows; on the 200M-row query 8 x 4 x 8 with
  // 3072 rows measured 1.19 ms against 1.31 for the tile kernel)
  int chunk_rows = 0, chunks = 0, cap = 0;
  if (staged) {
    jf->staged_waves = (int)tune_get("filter.staged_waves", 8);
    chunk_rows = (int)tune_get("filter.chunk_rows", 4);
    chunks = (int)tune_get("filter.chunks", 8);
    cap = (int)(tune_get("filter.stage_rows", 3072) & ~int64_t(7));
    if (chunk_rows * jf->staged_waves > 64)
      return fail(OXR_EPLAN, "rocm.jit", "filter.chunk_rows x filter.staged_waves must be at most 64");
    if (chunk_rows * chunks > kMaxRowsPerLane)
      return fail(OXR_EPLAN, "rocm.jit", "filter.chunk_rows x filter.chunks must be at most 32");
    if ((int64_t)cap * stage_bytes > 128 * 1024)
      return fail(OXR_EPLAN, "rocm.jit",
                  "filter.stage_rows x " + std::to_string(stage_bytes) + " projected bytes per row exceeds 128 KiB of LDS");
    jf->staged = true;
    jf->staged_rows = chunk_rows * chunks;
    jf->stage_rows = cap;
  }
  std::string src;
  src += "// generated filter + projection kernel\n";
  src += "#include \"filter_jit.h\"\nusing namespace oxr;\nnamespace {\nstruct Q {\n";
  src += "  static constexpr int kValidOuts = " + std::to_string(nvalid) + ";\n";
  src += "  static constexpr int kVecOuts = " + std::to_string(nvec) + ";\n";
  src += "  struct In {\n" + in_fields + "  };\n";
  src += "  __device__ static inline In ins(const JitArgs &p, uint64_t base) {\n    In i;\n" + in_init +
         "    return i;\n  }\n";
  src += "  struct Out {\n" + out_fields + "  };\n";
  src += "  __device__ static inline Out outs(const JitArgs &p, uint64_t base, uint64_t at) {\n    Out o;\n" +
         out_init + "    return o;\n  }\n";
  if (staged) {
    src += "  static constexpr int kStageBytes = " + std::to_string(stage_bytes) + ";\n";
    src += "  __device__ static inline Out outs_stage(char *stage, uint32_t cap) {\n    Out o;\n" + stage_init +
           "    (void)stage;\n    (void)cap;\n    return o;\n  }\n";
    src += "  __device__ static inline void copy_out(const Out &o, const char *stage, uint32_t cap, uint32_t n, "
           "int tid, int nt) {\n" +
           copy_code + "    (void)o;\n    (void)stage;\n    (void)cap;\n    (void)n;\n    (void)tid;\n    (void)nt;\n  }\n";
  }
  src += row;
  src += "  __device__ static inline void load(const In &i, uint32_t j, bool in, Row &x) {\n" + load +
         "  }\n";
  src += "  __device__ static inline bool pred(const uint64_t *L, const Row &x, uint32_t j, bool in, "
         "bool &err) {\n" + pred_code + "  }\n";
  src += "  __device__ static inline void project(const uint64_t *L, const Out &o, const Row &x, uint32_t j, "
         "uint32_t q, uint32_t bitpos, uint32_t *bits, int stride, bool &err) {\n" +
         proj_code + "  }\n";
  src += "  __device__ static inline void vectors(const Out &o, bool sel, uint32_t j, uint32_t q, "
         "int lane) {\n" +
         vec_code + "  }\n";
  src += "};\n}  // namespace\n";
  // the staged kernel is a module of its own, compiled when a launch first
  // takes it (small inputs and high kept fractions never do)
  if (staged)
    jf->staged_source = src + "extern \"C\" __global__ __launch_bounds__(" + std::to_string(64 * jf->staged_waves) +
                        ") void oxr_filter_staged(const JitArgs p) {\n  filter_staged<Q, " +
                        std::to_string(chunk_rows) + ", " + std::to_string(chunks) + ", " +
                        std::to_string(jf->staged_waves) + ", " + std::to_string(cap) + ">(p);\n}\n";
  src += "extern \"C\" __global__ __launch_bounds__(" + std::to_string(64 * jf->waves) +
         ") void oxr_filter_jit(const JitArgs p) {\n";
  src += std::string("  ") + (jf->persistent ? "filter_persistent" : "filter_tile") + "<Q, " +
         std::to_string(jf->rows_per_lane) + ", " + std::to_string(jf->waves) + ">(p);\n}\n";
  // one workgroup per batch (many-batch inputs, knob filter.wg): two
  // register sets of wg_rows rows per lane, wg_waves waves (knobs
  // filter.wg_rows, filter.wg_waves)
  jf->wg_rows = (int)tune_get("filter.wg_rows", 8);
  jf->wg_waves = (int)tune_get("filter.wg_waves", 8);
  jf->has_wg = tune_get("filter.wg", 0) > 0;  // emitted only when it may run (hipRTC time)
  const int wg_per_simd = jf->wg_waves / 4;
  if (jf->has_wg)
    src += "extern \"C\" __global__ __launch_bounds__(" + std::to_string(64 * jf->wg_waves) +
         ") __attribute__((amdgpu_waves_per_eu(" + std::to_string(wg_per_simd) + ", " +
         std::to_string(wg_per_simd) + "))) void oxr_filter_wg(const JitArgs p) {\n  filter_batch_wg<Q, " +
         std::to_string(jf->wg_rows) + ", " + std::to_string(jf->wg_waves) + ">(p);\n}\n";
  jf->source = src;
  return OXR_OK;
}

struct Entry {
  hipModule_t mod = nullptr;
  std::map<std::string, hipFunction_t> fns;  // kernels looked up so far
};

std::mutex g_jit_mu;
std::unordered_map<std::string, Entry> g_jit_cache;

// Compiles generated source for `arch` into a code object.
int rtc_compile(const std::string &arch, const std::string &src, std::vector<char> *code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "oxr_filter_jit.hip", kJitHeaderCount, kJitHeaderText,
                          kJitHeaderNames) != HIPRTC_SUCCESS)
    return fail(OXR_EDEVICE, "rocm.jit", "hiprtcCreateProgram failed");
  const std::string arch_opt = "--offload-arch=" + arch;
  const char *opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(OXR_EEXECUTION, "rocm.jit",
                std::string("generated kernel did not compile (") + hiprtcGetErrorString(rc) + "):\n" + log +
                    "\nsource:\n" + src);
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  code->assign(size, 0);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return OXR_OK;
}

// The code object for `src` on `dev`, compiled once per process and device.
// With `cached_only`, a source not compiled before is OXR_EUNSUPPORTED (no
// error recorded): nothing is compiled or loaded.
int compile(oxr_device *dev, const std::string &src, const char *kernel, hipFunction_t *fn,
            bool cached_only = false) {
  const std::string arch = dev->arch.substr(0, dev->arch.find(':'));
  const std::string key = arch + "/" + std::to_string(dev->ordinal) + "\n" + src;
  std::lock_guard<std::mutex> lock(g_jit_mu);
  auto it = g_jit_cache.find(key);
  if (it != g_jit_cache.end()) {
    auto f = it->second.fns.find(kernel);
    if (f != it->second.fns.end()) {
      *fn = f->second;
      return OXR_OK;
    }
    hipFunction_t h = nullptr;
    const hipError_t e = hipModuleGetFunction(&h, it->second.mod, kernel);
    if (e != hipSuccess) return fail_hip(e, "looking up a generated kernel");
    it->second.fns.emplace(kernel, h);
    *fn = h;
    return OXR_OK;
  }
  if (cached_only) return OXR_EUNSUPPORTED;
  std::vector<char> code;
  OXR_TRY(rtc_compile(arch, src, &code));
  int prev = 0;
  OXR_HIP(hipGetDevice(&prev));
  if (prev != dev->ordinal) OXR_HIP(hipSetDevice(dev->ordinal));
  Entry ent;
  hipFunction_t h = nullptr;
  hipError_t e = hipModuleLoadData(&ent.mod, code.data());
  if (e == hipSuccess) e = hipModuleGetFunction(&h, ent.mod, kernel);
  if (prev != dev->ordinal) hipSetDevice(prev);
  if (e != hipSuccess) return fail_hip(e, "loading a generated kernel");
  ent.fns.emplace(kernel, h);
  g_jit_cache.emplace(key, ent);
  *fn = h;
  return OXR_OK;
}

// The aggregation kernel of one shape: filter, key (a column, one global
// group, or the group a join gives the probe key) and each aggregate's
// contribution, over the aggregate_tiles skeleton (aggregate_jit.h).
int generate_aggregate(const std::vector<DevCol> &cols, const AggJitShape &sh, JitAgg *ja) {
  const int na = (int)sh.kinds.size();
  if (na < 1 || na > kMaxAggs || cols.size() > (size_t)kAggCols) return OXR_EUNSUPPORTED;
  std::vector<char> val_used(cols.size(), 0), valid_used(cols.size(), 0);
  JitFilter scratch;
  Gen g{cols, &scratch, val_used, valid_used};
  if (tune_get("jit.params", 0) > 0 || sh.multi > 0) g.params = &ja->lits;
  std::string pred_code = "    return true;\n";
  if (sh.filter && sh.filter->count > 0) {
    g.e = sh.filter;
    g.code.clear();
    const std::string res = g.boolean(0, sh.filter->count);
    pred_code = g.code + "    return " + res + ";\n";
  }
  // the key
  std::string key_code;
  auto column_ok = [&](int c) -> std::string {
    val_used[c] = 1;
    if (!cols[c].validity) return "in";
    valid_used[c] = 1;
    return "(in && row_bit(x.w" + std::to_string(c) + ", j))";
  };
  if (sh.join) {
    const int c = sh.key_col;  // the probe key
    const std::string ok = column_ok(c);
    key_code = "    if (!" + ok + ") return false;\n    return join_group(p, (int64_t)x.c" + std::to_string(c) +
               ", key, kv, hid);\n";
  } else if (sh.key_col >= 0) {
    const int c = sh.key_col;
    const std::string ok = column_ok(c);
    key_code = "    key = (int64_t)x.c" + std::to_string(c) + ";\n    kv = " + ok + ";\n    return true;\n";
  } else {
    key_code = "    key = 0;\n    kv = false;  // one global group: the NULL-key slot\n    return true;\n";
  }
  // the contributions
  std::string contrib = "    const bool in = sel;\n    (void)in;\n";
  std::string kinds, nulls, swords;
  bool any_null = false;
  int spill_words = 1;  // the group index, then the mask and the contributions
  for (int a = 0; a < na; a++) {
    any_null |= sh.nullable[a] != 0;
    const std::string as = std::to_string(a);
    const uint8_t k = sh.kinds[a];
    kinds += (a ? ", " : "") + std::to_string(k);
    nulls += (a ? ", " : "") + std::string(sh.nullable[a] ? "1" : "0");
    if (k == A_COUNT_STAR) {
      contrib += "    c[" + as + "] = 1;\n    cm |= (uint32_t)sel << " + as + ";\n";
      swords += (a ? ", 0" : "0");
      continue;
    }
    g.e = sh.inputs[a];
    g.code.clear();
    const Val v = g.value(0, sh.inputs[a]->count);
    // the 32-bit words of a spilled contribution (SpillRec)
    const bool narrow = ((k == A_SUM_I64 || k == A_MIN_I64 || k == A_MAX_I64) && v.t == OXR_INT32) ||
                        k == A_MIN_F32 || k == A_MAX_F32;
    const int w = k == A_COUNT ? 0 : narrow ? 1 : 2;
    swords += (a ? ", " : "") + std::to_string(w);
    spill_words += w;
    std::string cv;
    switch (k) {
      case A_COUNT: cv = "1"; break;
      case A_SUM_I64:
      case A_MIN_I64:
      case A_MAX_I64: cv = "(int64_t)" + v.v; break;
      case A_SUM_F64: cv = "(int64_t)f64_bits((double)" + v.v + ")"; break;
      case A_MIN_F64:
      case A_MAX_F64: cv = "total_order64(" + v.v + ")"; break;
      default: cv = "total_order32(" + v.v + ")"; break;
    }
    contrib += "    {\n" + g.code + "    c[" + as + "] = " + cv + ";\n    cm |= (uint32_t)(sel && " + v.ok +
               ") << " + as + ";\n    }\n";
  }
  if (g.unsupported) return OXR_EUNSUPPORTED;
  std::string in_fields, in_init, row, load;
  int vgprs = 0;
  row_code(cols, val_used, valid_used, &in_fields, &in_init, &row, &load, &vgprs);
  ja->rows_per_lane = pick_rows(std::max(vgprs + 2 * na, 1), "agg.rows");
  const int la = na <= kLdsAggs ? na : 0;
  std::string src;
  src += "// generated aggregation kernel\n";
  src += "#include \"filter_jit.h\"\n#include \"aggregate_jit.h\"\nusing namespace oxr;\nnamespace {\nstruct Q {\n";
  src += "  static constexpr int kAggs = " + std::to_string(na) + ";\n";
  src += "  static constexpr int kLdsAggs = " + std::to_string(la) + ";\n";
  src += "  static constexpr uint8_t kKind[kAggs] = {" + kinds + "};\n";
  src += "  static constexpr uint8_t kNullable[kAggs] = {" + nulls + "};\n";
  src += "  static constexpr uint8_t kSpillWords[kAggs] = {" + swords + "};\n";
  src += std::string("  static constexpr bool kAnyNullable = ") + (any_null ? "true" : "false") + ";\n";
  src += std::string("  static constexpr bool kExactF64 = ") + (tune_get("agg.exact_f64", 0) > 0 ? "true" : "false") +
         ";\n";
  // prepared mode: the literal count, and whether an aggregate input reads one
  src += "  static constexpr int kLits = " + std::to_string(ja->lits.size()) + ";\n";
  // the spill's records are read once by the scatter and once by the fold
  // (knob agg.spill_nt, default on: nontemporal loads; on the 500M x 10M
  // star the fold 1.34 -> 1.29-1.31 ms, the scatter 1.16 -> 1.12-1.15)
  src += std::string("  static constexpr bool kSpillNT = ") + (tune_get("agg.spill_nt", 1) > 0 ? "true" : "false") +
         ";\n";
  src += std::string("  static constexpr bool kContribLits = ") +
         (contrib.find("L[") != std::string::npos ? "true" : "false") + ";\n";
  src += "  struct In {\n" + in_fields + "  };\n";
  src += "  __device__ static inline In ins(const AggJitArgs &p, uint64_t base) {\n    In i;\n" + in_init +
         "    return i;\n  }\n";
  src += row;
  src += "  __device__ static inline void load(const In &i, uint32_t j, bool in, Row &x) {\n" + load + "  }\n";
  src += "  __device__ static inline bool pred(const uint64_t *L, const Row &x, uint32_t j, bool in, "
         "bool &err) {\n" + pred_code + "  }\n";
  src += "  __device__ static inline bool key(const AggJitArgs &p, const Row &x, uint32_t j, bool in, "
         "int64_t &key, bool &kv, int32_t &hid) {\n    hid = -2;\n" + key_code + "  }\n";
  src += "  __device__ static inline void contrib(const uint64_t *L, const Row &x, uint32_t j, bool sel, "
         "int64_t (&c)[kAggs], uint32_t &cm, bool &err) {\n" + contrib + "  }\n";
  src += "};\n}  // namespace\n";
  if (sh.multi > 0) {
    // batched requests: rows per lane from the knob agg.multi_rows
    const int64_t mr = tune_get("agg.multi_rows", 4);
    ja->multi_rows = (int)(mr > 0 ? mr : 4);
    for (int i = 0; i < 4; i++) {
      if (kMultiK[i] > sh.multi) break;
      const std::string k = std::to_string(kMultiK[i]);
      src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_multi" + k +
             "(const AggJitArgs p) {\n  aggregate_multi<Q, " + std::to_string(ja->multi_rows) + ", " + k +
             ">(p);\n}\n";
      src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_multi_final" + k +
             "(const AggJitArgs p) {\n  aggregate_multi_final<Q, " + k + ">(p);\n}\n";
    }
    ja->source = src;
    return OXR_OK;
  }
  src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_jit(const AggJitArgs p) {\n";
  src += "  aggregate_tiles<Q, " + std::to_string(ja->rows_per_lane) + ">(p);\n}\n";
  // high-cardinality skewed keys: the hot-group kernel (aggregate_hot), for
  // shapes whose aggregates all fit the block's LDS and sum FP64 plainly
  // the spill's scatter and fold for this shape's records: shapes whose
  // rows may spill (a key, every aggregate in the block table, plain FP64
  // sums)
  ja->spill = la == na && tune_get("agg.exact_f64", 0) == 0 && sh.key_col != -1;
  if (ja->spill)
    src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_spill_scatter(const SpillArgs a) {\n"
           "  spill_scatter<Q>(a);\n}\n"
           "extern \"C\" __global__ __launch_bounds__(512) void oxr_spill_fold(const SpillArgs a) {\n"
           "  spill_fold<Q>(a);\n}\n";
  ja->spill_words = spill_words + (any_null ? 1 : 0);
  ja->hot = ja->spill;
  if (ja->hot) {
    // the shape's rows per lane (two tiles of them in registers, the next
    // tile's loads in flight), or the knob agg.hot_lane_rows
    const int64_t hs = tune_get("agg.hot_lane_rows", 0);
    const int hot_rows = hs > 0 ? (int)hs : ja->rows_per_lane;
    src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_hot(const AggJitArgs p) {\n"
           "  aggregate_hot<Q, " + std::to_string(hot_rows) + ">(p);\n}\n";
  }
  // a global aggregate: per-lane register accumulators (aggregate_keyless)
  ja->keyless = sh.key_col == -1 && !sh.join;
  if (ja->keyless)
    src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_keyless(const AggJitArgs p) {\n"
           "  aggregate_keyless<Q, " + std::to_string(ja->rows_per_lane) + ">(p);\n}\n";
  ja->dense = sh.key_col >= 0 && !sh.join;
  if (ja->dense) {
    // small key ranges: per-lane LDS accumulators (aggregate_dense)
    for (int ds : {16, 32})
      src += "extern \"C\" __global__ __launch_bounds__(256) void oxr_aggregate_dense" + std::to_string(ds) +
             "(const AggJitArgs p) {\n  aggregate_dense<Q, " + std::to_string(ja->rows_per_lane) + ", " +
             std::to_string(ds) + ">(p);\n}\n";
  }
  ja->source = src;
  return OXR_OK;
}

}  // namespace

int jit_aggregate(oxr_device *dev, const std::vector<DevCol> &cols, const AggJitShape &shape, JitAgg *out) {
  JitAgg ja;
  const int rc = generate_aggregate(cols, shape, &ja);
  if (rc != OXR_OK) return rc;
  if (shape.multi > 0) {
    for (int i = 0; i < 4 && kMultiK[i] <= shape.multi; i++) {
      const std::string k = std::to_string(kMultiK[i]);
      OXR_TRY(compile(dev, ja.source, ("oxr_aggregate_multi" + k).c_str(), &ja.multi_fn[i]));
      OXR_TRY(compile(dev, ja.source, ("oxr_aggregate_multi_final" + k).c_str(), &ja.multi_final[i]));
    }
    *out = std::move(ja);
    return OXR_OK;
  }
  OXR_TRY(compile(dev, ja.source, "oxr_aggregate_jit", &ja.fn));
  if (ja.dense) {
    OXR_TRY(compile(dev, ja.source, "oxr_aggregate_dense16", &ja.dense16));
    OXR_TRY(compile(dev, ja.source, "oxr_aggregate_dense32", &ja.dense32));
  }
  if (ja.hot) OXR_TRY(compile(dev, ja.source, "oxr_aggregate_hot", &ja.hot_fn));
  if (ja.keyless) OXR_TRY(compile(dev, ja.source, "oxr_aggregate_keyless", &ja.keyless_fn));
  if (ja.spill) {
    OXR_TRY(compile(dev, ja.source, "oxr_spill_scatter", &ja.scatter_fn));
    OXR_TRY(compile(dev, ja.source, "oxr_spill_fold", &ja.fold_fn));
  }
  *out = std::move(ja);
  return OXR_OK;
}

int jit_aggregate_generate(const std::vector<DevCol> &cols, const AggJitShape &shape, JitAgg *out) {
  return generate_aggregate(cols, shape, out);
}

int jit_aggregate_source(const std::vector<DevCol> &cols, const AggJitShape &shape, bool do_compile,
                         std::string *src) {
  JitAgg ja;
  const int rc = generate_aggregate(cols, shape, &ja);
  if (rc != OXR_OK) return rc;
  if (do_compile) {
    std::vector<char> code;
    OXR_TRY(rtc_compile("gfx950", ja.source, &code));
  }
  *src = ja.source;
  return OXR_OK;
}

namespace {
}  // namespace

namespace {
// Generated filter kernels by query shape: the columns' types, dims and
// nullability, the expression nodes (literals included) and the knob epoch.
// Generating the source takes ~25 us a call, a sizeable part of a 1 ms launch.
std::mutex g_gen_mu;
std::unordered_map<std::string, JitFilter> g_gen;

std::string shape_key(const std::vector<DevCol> &cols, const oxr_expr *pred, const oxr_expr *proj,
                      int32_t nproj) {
  std::string key;
  auto put = [&](const void *p, size_t n) { key.append((const char *)p, n); };
  const uint64_t epoch = tune_epoch();
  put(&epoch, sizeof epoch);
  const int32_t nc = (int32_t)cols.size();
  put(&nc, sizeof nc);
  for (const DevCol &c : cols) {
    const int32_t v[3] = {c.type, c.dim, c.validity ? 1 : 0};
    put(v, sizeof v);
  }
  auto put_expr = [&](const oxr_expr *e) {
    const int32_t n = e ? e->count : -1;
    put(&n, sizeof n);
    if (e && e->count > 0) put(e->nodes, (size_t)e->count * sizeof(oxr_expr_node));
  };
  put_expr(pred);
  put(&nproj, sizeof nproj);
  for (int32_t k = 0; k < nproj; k++) put_expr(&proj[k]);
  return key;
}
}  // namespace

int jit_filter(oxr_device *dev, const std::vector<DevCol> &cols, const oxr_expr *pred,
               const oxr_expr *proj, int32_t nproj, const std::vector<ValueProg> &vps,
               JitFilter *out, bool cached_only) {
  JitFilter jf;
  const std::string key = shape_key(cols, pred, proj, nproj);
  bool hit = false;
  {
    std::lock_guard<std::mutex> lock(g_gen_mu);
    auto it = g_gen.find(key);
    if (it != g_gen.end()) {
      jf = it->second;
      hit = true;
    }
  }
  if (!hit) {
    const int rc = generate(cols, pred, proj, nproj, vps, &jf);
    if (rc != OXR_OK) return rc;
    std::lock_guard<std::mutex> lock(g_gen_mu);
    if (g_gen.size() >= 4096) g_gen.clear();
    g_gen.emplace(key, jf);
  }
  OXR_TRY(compile(dev, jf.source, "oxr_filter_jit", &jf.fn, cached_only));
  if (jf.has_wg) OXR_TRY(compile(dev, jf.source, "oxr_filter_wg", &jf.wg_fn, cached_only));
  *out = std::move(jf);
  return OXR_OK;
}

namespace {
std::mutex g_kept_mu;
std::unordered_map<hipFunction_t, double> g_kept;  // staged kernel -> kept fraction of its last launch
}  // namespace

double jit_kept_fraction(hipFunction_t fn) {
  std::lock_guard<std::mutex> lock(g_kept_mu);
  auto it = g_kept.find(fn);
  return it == g_kept.end() ? -1.0 : it->second;
}

void jit_note_kept(hipFunction_t fn, double fraction) {
  std::lock_guard<std::mutex> lock(g_kept_mu);
  g_kept[fn] = fraction;
}

int jit_filter_staged(oxr_device *dev, JitFilter *jf, bool cached_only) {
  if (jf->staged_fn) return OXR_OK;
  if (!jf->staged) return OXR_EUNSUPPORTED;
  return compile(dev, jf->staged_source, "oxr_filter_staged", &jf->staged_fn, cached_only);
}

int jit_launch_raw(hipFunction_t fn, unsigned grid, hipStream_t s, void *args, size_t size, unsigned block) {
  void *cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                 HIP_LAUNCH_PARAM_END};
  OXR_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, s, nullptr, cfg));
  return OXR_OK;
}

int jit_launch(hipFunction_t fn, unsigned grid, hipStream_t s, const JitArgs &args, unsigned block) {
  JitArgs a = args;
  return jit_launch_raw(fn, grid, s, &a, sizeof a, block);
}

}  // namespace oxr

using namespace oxr;

extern "C" int oxr_filter_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_expr *predicate,
                                        const oxr_expr *projection, int32_t nproj, int32_t compile,
                                        char *buf, size_t cap, size_t *len) {
  clear_error();
  if ((ncols > 0 && !cols) || ncols < 0 || (nproj > 0 && !projection) || nproj < 0)
    return fail(OXR_EPLAN, "rocm.jit", "null argument");
  std::vector<DevCol> dc(ncols);
  for (int i = 0; i < ncols; i++) {
    dc[i].type = cols[i].type;
    dc[i].dim = cols[i].dim;
    dc[i].len = cols[i].len;
    dc[i].validity = cols[i].validity;  // only NULL or not matters
  }
  Program prog;
  OXR_TRY(compile_predicate(predicate, dc, &prog));
  std::vector<ValueProg> vps(nproj);
  for (int i = 0; i < nproj; i++) OXR_TRY(compile_value(&projection[i], dc, &prog, &vps[i]));
  JitFilter jf;
  const int rc = generate(dc, predicate, projection, nproj, vps, &jf);
  if (rc == OXR_EPLAN) return rc;  // a knob asks for a shape the kernels do not take
  if (rc != OXR_OK)
    return fail(OXR_EUNSUPPORTED, "rocm.jit", "outside what generated kernels cover (the interpreter runs it)");
  // the tile kernel's module, then the staged kernel's (a module of its own)
  if (compile) {
    std::vector<char> code;
    OXR_TRY(rtc_compile("gfx950", jf.source, &code));
    if (jf.staged) OXR_TRY(rtc_compile("gfx950", jf.staged_source, &code));
  }
  jf.source += jf.staged_source;
  if (len) *len = jf.source.size();
  if (buf && cap) {
    const size_t n = std::min(cap - 1, jf.source.size());
    std::memcpy(buf, jf.source.data(), n);
    buf[n] = 0;
  }
  return OXR_OK;
}

// Grouped hash aggregation: SUM / COUNT / MIN / MAX / COUNT(*) by one key.
//
// Replaces `ops::aggregate` (crates/oxidelake-device/src/cuda/ops.rs:623-1161)
// and its kernels (kernels/cuda/aggregation.cu), which size a table by the
// batch's rows, insert keys in one launch, number groups in another, run one
// launch per aggregate function (each re-probing the table) and merge the
// per-batch partial on the host. Here:
//  - the group table lives in HBM across batches (streaming state), open
//    addressing with a single 64-bit CAS on the key (INT64_MIN marks an empty
//    slot; the NULL key and the key INT64_MIN itself have dedicated slots), so
//    no claimed/published spin protocol is needed;
//  - one kernel per batch evaluates the optional fused filter, the key and
//    every aggregate's input expression, and folds each row into an
//    LDS-privatised block table (hot keys aggregate in LDS), which is flushed
//    to the HBM table with one atomic per (group, aggregate) per flush;
//  - SUM over integers wraps in Int64; SUM over floats is an FP64 atomic sum
//    from +0.0 (order differs from the CPU's sequential fold: the parity
//    contract is 1e-12 relative); MIN/MAX of floats are integer min/max over
//    totalOrder keys (kernels.rs:820-838).
// Semantics: crates/oxidelake-device/src/cpu/kernels.rs:478-818.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <memory>

#include "common.h"
#include "expr.h"
#include "join.h"
#include "aggregate_jit.h"
#include "jit.h"

namespace oxr {
namespace {

constexpr int V = 1;  // rows per lane (register use grows with it)
constexpr int kBlock = 256;
constexpr int kTileRows = kBlock * V;

struct AggDesc {
  uint8_t kind;
  uint8_t in_type;   // value type of the input expression
  uint8_t nullable;  // input can be NULL (then `has` flags are kept)
  uint8_t pad;
  ValueProg vp;      // the input: a slot or a program
};

struct Params {
  uint64_t n;
  DCol cols[kMaxCols];
  Ins ins[kMaxIns];
  Leaf leaves[kMaxLeaves];
  uint8_t bprog[kMaxBool];
  DivMagic magic[kMaxMagic];
  int8_t slot_cols[kMaxSlots];
  int32_t nslots;
  int32_t nb;
  int32_t key_slot;  // -1: one global group
  int32_t naggs;
  int32_t lds_aggs;  // aggregates folded in LDS (0 = all global)
  AggDesc aggs[kMaxAggs];
  Table t;
  uint32_t *error;   // division by zero
  // fused inner join: the group key of a probe row is the build side's
  // group column at each build row its key matches
  int32_t join;           // 0: the key is key_slot of the row itself
  int32_t probe_slot;     // join: slot of the probe key column
  JTable jt;
  DCol build_group;
  int32_t exact_f64;      // compensated FP64 sums (aggregate_jit.h)
};

template <int NS>
__global__ __launch_bounds__(kBlock) void aggregate_kernel(const Params p) {
  __shared__ int64_t s_keys[kLdsSlots + 2];
  __shared__ int64_t s_acc[kLdsAggs][kLdsSlots + 2];
  __shared__ int64_t s_lo[kLdsAggs][kLdsSlots + 2];  // FP64 sums: rounding errors
  __shared__ uint8_t s_has[kLdsAggs][kLdsSlots + 2];
  __shared__ uint32_t s_used;
  __shared__ uint32_t s_flush;

  const int tid = threadIdx.x;
  const int la = p.lds_aggs;
  for (int i = tid; i < kLdsSlots + 2; i += kBlock) {
    s_keys[i] = kEmpty;
    for (int a = 0; a < la; a++) {
      s_acc[a][i] = acc_init(p.aggs[a].kind);
      s_lo[a][i] = 0;
      s_has[a][i] = 0;
    }
  }
  if (tid == 0) {
    s_used = 0;
    s_flush = 0;
  }
  __syncthreads();

  bool div0 = false, ok = true;
  const uint64_t ntiles = (p.n + kTileRows - 1) / kTileRows;

  auto flush = [&]() {
    // every used LDS slot -> one global update per aggregate, then clear
    for (int i = tid; i < kLdsSlots + 2; i += kBlock) {
      const int64_t k = s_keys[i];
      const bool used = i < kLdsSlots ? k != kEmpty : k != kEmpty;  // specials carry a marker
      if (used) {
        uint64_t gs;
        if (i == kLdsSlots) gs = slot_of(p.t, false, 0, &ok);
        else if (i == kLdsSlots + 1) gs = slot_of(p.t, true, kEmpty, &ok);
        else gs = global_slot(p.t, k, &ok);
        if (ok) {
          for (int a = 0; a < la; a++) {
            const AggDesc d = p.aggs[a];
            const int64_t v = s_acc[a][i];
            if (p.exact_f64 && d.kind == A_SUM_F64)
              sum_f64(&p.t.acc[a][gs], &p.t.lo[a][gs], as_f64((uint64_t)v), as_f64((uint64_t)s_lo[a][i]));
            else if (v != acc_init(d.kind) || d.kind == A_SUM_F64)
              acc_apply_global(d.kind, &p.t.acc[a][gs], v);
            if (p.t.has[a] && s_has[a][i]) p.t.has[a][gs] = 1;
          }
        }
      }
      s_keys[i] = kEmpty;
      for (int a = 0; a < la; a++) {
        s_acc[a][i] = acc_init(p.aggs[a].kind);
        s_lo[a][i] = 0;
        s_has[a][i] = 0;
      }
    }
  };

  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t row0 = tile * kTileRows + (uint64_t)tid * V;
    Slots<V, NS> S;
    load_slots<V, NS>(p.cols, p.slot_cols, p.nslots, row0, p.n, S);
    uint32_t sel = eval_predicate<V, NS>(p.leaves, p.bprog, p.nb, p.ins, p.cols, S, p.magic, row0,
                                         p.n, &div0);
    static_assert(V == 1, "one row per lane");
    // the row's matches: itself, or the build rows its key joins
    uint32_t nmatch = sel & 1u;
    uint32_t jstart = 0;
    if (p.join && nmatch) {
      nmatch = 0;
#pragma unroll
      for (int k = 0; k < NS; k++) {
        if (p.probe_slot == k && (S.m[k] & 1u)) {
          const uint64_t js = join_find(p.jt, (int64_t)S.v[k][0]);
          if (js != UINT64_MAX) {
            jstart = p.jt.start[js];
            nmatch = p.jt.count[js];
          }
        }
      }
    }
    // the key of a plain row (a NULL key is a group of its own: validity
    // only picks the slot; one global group takes the NULL slot)
    int64_t rkey = 0;
    bool rkvalid = false;
    if (!p.join && p.key_slot >= 0) {
#pragma unroll
      for (int k = 0; k < NS; k++) {
        if (p.key_slot == k) {
          rkey = (int64_t)S.v[k][0];
          rkvalid = S.m[k] & 1u;
        }
      }
    }
    // aggregate inputs, once per row
    uint64_t av[kMaxAggs];
    uint32_t avm = 0;
    if (nmatch) {
#pragma unroll
      for (int a = 0; a < kMaxAggs; a++) {
        if (a >= p.naggs) break;
        const AggDesc d = p.aggs[a];
        uint64_t v[V] = {0};
        uint32_t vm = 1;
        if (d.kind != A_COUNT_STAR) eval_value<V, NS>(d.vp, p.ins, S, p.magic, v, vm, &div0);
        av[a] = v[0];
        avm |= (vm & 1u) << a;
      }
    }
    for (uint32_t mi = 0; mi < nmatch; mi++) {
      int64_t k = rkey;
      bool kv = rkvalid;
      if (p.join) {
        const uint64_t b = p.jt.rows[jstart + mi];
        const DCol &g = p.build_group;
        kv = !g.valid || ((g.valid[b >> 5] >> (b & 31)) & 1u);
        k = g.type == OXR_INT32 ? (int64_t)((const int32_t *)g.values)[b] : ((const int64_t *)g.values)[b];
      }
      // the block table's slot (-1: go global)
      int lslot = -1;
      if (la) {
        if (!kv) {
          lslot = kLdsSlots;
          s_keys[kLdsSlots] = 0;  // marks the NULL-key slot used
        } else if (k == kEmpty) {
          lslot = kLdsSlots + 1;
          s_keys[kLdsSlots + 1] = 0;
        } else {
          uint32_t h = lds_hash(k);
          for (int pr = 0; pr < kLdsProbe; pr++) {
            const int64_t cur = s_keys[h];
            if (cur == k) {
              lslot = (int)h;
              break;
            }
            if (cur == kEmpty) {
              const unsigned long long old = atomicCAS((unsigned long long *)&s_keys[h],
                                                       (unsigned long long)kEmpty, (unsigned long long)k);
              if ((int64_t)old == kEmpty) {
                atomicAdd(&s_used, 1u);
                lslot = (int)h;
                break;
              }
              if ((int64_t)old == k) {
                lslot = (int)h;
                break;
              }
            }
            h = (h + 1) & (kLdsSlots - 1);
          }
        }
      }
      // the HBM slot, for rows the block table could not take (or for
      // aggregates past the ones it holds)
      uint64_t gslot = 0;
      if (lslot < 0 || la < p.naggs) gslot = slot_of(p.t, kv, k, &ok);
#pragma unroll
      for (int a = 0; a < kMaxAggs; a++) {
        if (a >= p.naggs) break;
        const AggDesc d = p.aggs[a];
        if (!((avm >> a) & 1u) && d.kind != A_COUNT_STAR) continue;
        const int64_t c = contribution(d.kind, d.in_type, av[a]);
        if (a < la && lslot >= 0) {
          if (p.exact_f64 && d.kind == A_SUM_F64) sum_f64(&s_acc[a][lslot], &s_lo[a][lslot], as_f64((uint64_t)c), 0.0);
          else acc_apply_lds(d.kind, &s_acc[a][lslot], c);
          if (d.nullable) s_has[a][lslot] = 1;
        } else {
          if (!ok) continue;
          if (p.exact_f64 && d.kind == A_SUM_F64) sum_f64(&p.t.acc[a][gslot], &p.t.lo[a][gslot], as_f64((uint64_t)c), 0.0);
          else acc_apply_global(d.kind, &p.t.acc[a][gslot], c);
          if (p.t.has[a]) p.t.has[a][gslot] = 1;
        }
      }
    }
    if (la) {
      __syncthreads();
      if (s_used > kLdsSlots / 2) {
        flush();
        __syncthreads();
        if (tid == 0) s_used = 0;
        __syncthreads();
      }
    }
  }
  if (la) {
    __syncthreads();
    flush();
  }
  if (div0) atomicOr(p.error, 1u);
  if (!ok) atomicOr(p.t.overflow, 1u);
}

// ------------------------------------------------------------------ finish
// Compacts (key, slot) of the occupied ordinary slots; one counter atomic
// per block (same-address atomics serialise at ~10 ns each).
__global__ __launch_bounds__(256) void compact_kernel(Table t, int64_t *out_keys, uint64_t *out_slots,
                                                      uint64_t *count) {
  __shared__ uint32_t s_wave[4];
  __shared__ unsigned long long s_base;
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const int64_t k = i < t.cap ? t.keys[i] : kEmpty;
  const bool used = k != kEmpty;
  const uint64_t m = __ballot(used);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_wave[w] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    s_base = total ? atomicAdd((unsigned long long *)count, (unsigned long long)total) : 0;
  }
  __syncthreads();
  if (used) {
    uint64_t at = s_base + __popcll(m & ((1ull << lane) - 1));
    for (int v = 0; v < w; v++) at += s_wave[v];
    out_keys[at] = k;
    out_slots[at] = i;
  }
}

struct OutCol {
  void *values;
  uint32_t *valid;
  uint8_t kind, in_type, out_type, nullable;
};

// A validity bit per lane of the wave, written as whole words: lane 0 writes
// the word of lanes 0-31, lane 32 that of lanes 32-63 (j is 64-aligned per
// wave, so each word has one writer).
__device__ inline void put_valid_bits(uint32_t *valid, uint64_t j, uint64_t total, bool bit) {
  const uint64_t m = __ballot(bit);
  const int lane = threadIdx.x & 63;
  if ((lane == 0 || lane == 32) && j < total) valid[j >> 5] = (uint32_t)(lane == 0 ? m : m >> 32);
}

// Row j of the output: group slots[j]; optional special rows appended.
__global__ __launch_bounds__(256) void gather_kernel(Table t, const int64_t *keys_sorted,
                                                     const uint64_t *slots_sorted, uint64_t g_ordinary,
                                                     int has_min_key, int has_null, int key_type, void *out_key,
                                                     uint32_t *out_key_valid, OutCol *outs, int naggs,
                                                     uint64_t total, int empty_group) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const bool in = j < total;
  uint64_t slot = 0;
  int64_t key = 0;
  bool key_valid = in;
  if (in) {
    // order: the INT64_MIN key first (smallest), ordinary keys ascending, NULL last
    if (has_min_key && j == 0) {
      slot = t.cap + 1;
      key = INT64_MIN;
    } else if (has_null && j == total - 1) {
      slot = t.cap;
      key = 0;
      key_valid = false;
    } else {
      const uint64_t r = j - (has_min_key ? 1 : 0);
      slot = slots_sorted[r];
      key = keys_sorted[r];
    }
    if (out_key) {
      if (key_type == OXR_INT32) ((int32_t *)out_key)[j] = (int32_t)key;
      else ((int64_t *)out_key)[j] = key;
    }
  }
  if (out_key) put_valid_bits(out_key_valid, j, total, key_valid);
  for (int a = 0; a < naggs; a++) {
    const OutCol o = outs[a];
    bool valid = false;
    if (in) {
      int64_t v = t.acc[a][slot];
      if (o.kind == A_SUM_F64 && t.lo[a])
        v = (int64_t)f64_bits(sum_value(as_f64((uint64_t)v), as_f64((uint64_t)t.lo[a][slot])));
      if (o.kind == A_COUNT || o.kind == A_COUNT_STAR) valid = true;
      else if (t.has[a]) valid = t.has[a][slot] != 0;
      else valid = !empty_group;  // a non-nullable input: every row of a group holds a value
      uint64_t outv;
      switch (o.kind) {
        case A_MIN_F64:
        case A_MAX_F64: {
          int64_t b = v ^ (int64_t)(((uint64_t)(v >> 63)) >> 1);
          outv = (uint64_t)b;
          break;
        }
        case A_MIN_F32:
        case A_MAX_F32: {
          int32_t k32 = (int32_t)v;
          int32_t b = k32 ^ (int32_t)(((uint32_t)(k32 >> 31)) >> 1);
          outv = (uint64_t)(uint32_t)b;
          break;
        }
        default: outv = (uint64_t)v;
      }
      if (o.out_type == OXR_INT32 || o.out_type == OXR_FLOAT32) ((uint32_t *)o.values)[j] = (uint32_t)outv;
      else ((uint64_t *)o.values)[j] = outv;
    }
    if (o.valid) put_valid_bits(o.valid, j, total, valid);
  }
}

__global__ void fill64_kernel(int64_t *p, uint64_t n, int64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Re-inserts every occupied slot of `from` into the empty table `to`.
__global__ void rehash_kernel(Table from, Table to, int naggs) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= from.cap + 2) return;
  uint64_t dst;
  bool ok = true;
  if (i < from.cap) {
    const int64_t k = from.keys[i];
    if (k == kEmpty) return;
    dst = global_slot(to, k, &ok);
    if (!ok) {
      atomicOr(to.overflow, 1u);
      return;
    }
  } else {
    dst = i - from.cap + to.cap;  // special slots keep their role
  }
  for (int a = 0; a < naggs; a++) {
    to.acc[a][dst] = from.acc[a][i];
    if (from.lo[a]) to.lo[a][dst] = from.lo[a][i];
    if (from.has[a]) to.has[a][dst] = from.has[a][i];
  }
}

int fill64(oxr_stream *s, int64_t *p, uint64_t n, int64_t v) {
  if (!n) return OXR_OK;
  unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill64_kernel, dim3(blocks), dim3(256), 0, s->s, p, n, v);
  OXR_HIP(hipGetLastError());
  return OXR_OK;
}

// Key columns of more rows than this (knob agg.range_rows) get their value
// range measured before the table is sized (below it, sizing for a key per
// row costs less).
constexpr int64_t kRangeCheckRows = 1 << 16;

uint64_t next_pow2(uint64_t x) {
  uint64_t c = 16;
  while (c < x) c <<= 1;
  return c;
}

}  // namespace
}  // namespace oxr

using namespace oxr;

struct oxr_agg_state {
  oxr_device *dev = nullptr;
  oxr::StreamCore *owner = nullptr;
  int32_t key_col = -1;  // -1: one global group; -2: keys from a join's build side
  int32_t key_slot = -1;
  int32_t key_type = OXR_INT64;
  std::string key_name;
  uint64_t hint = 0;
  Program prog;  // filter + aggregate input programs
  std::vector<AggDesc> aggs;
  std::vector<int32_t> out_types;
  std::vector<std::string> names;
  std::vector<int32_t> col_types;  // schema the spec was compiled against
  Table t{};
  void *mem = nullptr;  // one arena block holding the table
  uint64_t groups = 0;  // distinct ordinary keys seen (after the last update)
  // an update failed after some of its rows were applied: the groups are no
  // longer those of the batches folded so far, and every later call fails
  bool poisoned = false;
  // the spec as expressions, for the generated kernels
  std::vector<oxr_expr_node> filter_nodes;
  std::vector<std::vector<oxr_expr_node>> input_nodes;
  std::vector<int32_t> funcs;
  // the group of every slot of the join table last folded through
  void *gos = nullptr;
  bool gos32 = true;
  const void *gos_table = nullptr, *gos_build = nullptr;
  int32_t gos_col = -1;
};

namespace oxr {
namespace {

int table_alloc(oxr_stream *s, oxr_agg_state *st, uint64_t cap, Table *out, void **mem) {
  const int na = (int)st->aggs.size();
  const uint64_t slots = cap + 2;
  size_t bytes = slots * 8 * (1 + na) + 64;
  for (int a = 0; a < na; a++) {
    if (st->aggs[a].nullable) bytes += slots * 4 + 16;
    if (st->aggs[a].kind == A_SUM_F64) bytes += slots * 8;
  }
  bytes += 64;  // seen, distinct, overflow
  void *m = nullptr;
  OXR_TRY(stream_alloc(s, bytes, &m));
  char *p = (char *)m;
  Table t{};
  t.cap = cap;
  t.keys = (int64_t *)p;
  p += slots * 8;
  for (int a = 0; a < na; a++) {
    t.acc[a] = (int64_t *)p;
    p += slots * 8;
  }
  for (int a = 0; a < na; a++) {
    t.lo[a] = nullptr;
    if (st->aggs[a].kind == A_SUM_F64) {
      t.lo[a] = (int64_t *)p;
      p += slots * 8;
    }
  }
  p = (char *)(((uintptr_t)p + 15) & ~(uintptr_t)15);
  for (int a = 0; a < na; a++) {
    t.has[a] = nullptr;
    if (st->aggs[a].nullable) {
      t.has[a] = (uint32_t *)p;
      p += slots * 4;
    }
  }
  p = (char *)(((uintptr_t)p + 15) & ~(uintptr_t)15);
  t.seen = (uint32_t *)p;
  t.distinct = (uint64_t *)(p + 16);
  t.overflow = (uint32_t *)(p + 24);
  int rc = fill64(s, t.keys, slots, kEmpty);
  for (int a = 0; a < na && rc == OXR_OK; a++) rc = fill64(s, t.acc[a], slots, acc_init(st->aggs[a].kind));
  for (int a = 0; a < na && rc == OXR_OK; a++)
    if (t.lo[a]) rc = fill64(s, t.lo[a], slots, 0);
  for (int a = 0; a < na && rc == OXR_OK; a++)
    if (t.has[a]) rc = fill_zero(s, t.has[a], slots * 4);
  if (rc == OXR_OK) rc = fill_zero(s, t.seen, 32);
  if (rc != OXR_OK) {
    stream_release(s, m);
    return rc;
  }
  *out = t;
  *mem = m;
  return OXR_OK;
}

// min and max of the valid values of an Int32/Int64 column, into mm[0..1]
// (set to INT64_MAX and INT64_MIN beforehand). Each thread reads 8
// consecutive values (one validity byte) per step; one atomic pair per block.
template <typename T, bool kValid>
__global__ __launch_bounds__(256) void key_range_kernel(const T *v, const uint32_t *valid, uint64_t n,
                                                        int64_t *mm) {
  __shared__ int64_t s_lo[4], s_hi[4];
  int64_t lo = kI64Max, hi = kI64Min;
  const uint64_t chunks = n / 8;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < chunks; c += (uint64_t)gridDim.x * blockDim.x) {
    T x[8];
    __builtin_memcpy(x, v + c * 8, sizeof x);
    const uint32_t m = kValid ? (valid[c >> 2] >> ((c & 3) * 8)) & 0xFFu : 0xFFu;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      if (!((m >> e) & 1u)) continue;
      const int64_t y = (int64_t)x[e];
      lo = y < lo ? y : lo;
      hi = y > hi ? y : hi;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n - chunks * 8) {  // the last n % 8 values
    const uint64_t i = chunks * 8 + threadIdx.x;
    if (!kValid || ((valid[i >> 5] >> (i & 31)) & 1u)) {
      const int64_t y = (int64_t)v[i];
      lo = y < lo ? y : lo;
      hi = y > hi ? y : hi;
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const int64_t ol = (int64_t)__shfl_xor((long long)lo, d, 64);
    const int64_t oh = (int64_t)__shfl_xor((long long)hi, d, 64);
    lo = ol < lo ? ol : lo;
    hi = oh > hi ? oh : hi;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_lo[w] = lo;
    s_hi[w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; i++) {
      lo = s_lo[i] < lo ? s_lo[i] : lo;
      hi = s_hi[i] > hi ? s_hi[i] : hi;
    }
    if (lo != kI64Max) atomicMin((long long *)&mm[0], (long long)lo);
    if (hi != kI64Min) atomicMax((long long *)&mm[1], (long long)hi);
  }
}

// The value range of a key column: its smallest valid value and how many
// distinct valid keys it can hold (max - min + 1; 0 when every value is
// NULL). One pass over the column.
struct KeyRange {
  bool known = false;
  int64_t min = 0;
  uint64_t count = 0;
};

int key_range(oxr_stream *s, const DevCol &c, uint64_t n, KeyRange *out) {
  void *m = nullptr;
  OXR_TRY(stream_alloc(s, 16, &m));
  int64_t *mm = (int64_t *)m;
  int rc = fill64(s, mm, 1, kI64Max);
  if (rc == OXR_OK) rc = fill64(s, mm + 1, 1, kI64Min);
  int64_t host[2] = {kI64Max, kI64Min};
  if (rc == OXR_OK) {
    const uint32_t *valid = (const uint32_t *)c.validity;
    const uint64_t grid = std::max<uint64_t>(
        1, std::min<uint64_t>((n / 8 + 255) / 256, (uint64_t)std::max(1, s->dev->cu_count) * 2));
    const dim3 g((unsigned)grid), b(256);
    {
      ProfScope prof(s->s, "key_range");
      if (c.type == OXR_INT32) {
        if (valid) hipLaunchKernelGGL((key_range_kernel<int32_t, true>), g, b, 0, s->s, (const int32_t *)c.values, valid, n, mm);
        else hipLaunchKernelGGL((key_range_kernel<int32_t, false>), g, b, 0, s->s, (const int32_t *)c.values, valid, n, mm);
      } else {
        if (valid) hipLaunchKernelGGL((key_range_kernel<int64_t, true>), g, b, 0, s->s, (const int64_t *)c.values, valid, n, mm);
        else hipLaunchKernelGGL((key_range_kernel<int64_t, false>), g, b, 0, s->s, (const int64_t *)c.values, valid, n, mm);
      }
    }
    hipError_t e = hipGetLastError();
    int64_t *pin = nullptr;
    if (e == hipSuccess) rc = stream_pinned(s, 16, (void **)&pin);
    if (rc == OXR_OK) {
      if (e == hipSuccess) e = hipMemcpyAsync(pin, mm, 16, hipMemcpyDeviceToHost, s->s);
      if (e == hipSuccess) e = hipStreamSynchronize(s->s);
      if (e != hipSuccess) rc = fail_hip(e, "key range");
      else std::memcpy(host, pin, 16);
    }
  }
  stream_release(s, m);
  if (rc != OXR_OK) return rc;
  out->known = true;
  out->min = host[0];
  out->count = host[1] < host[0] ? 0 : (uint64_t)host[1] - (uint64_t)host[0] + 1;  // 0 only for the full range
  if (out->count == 0 && host[1] >= host[0]) out->count = ~0ull;
  return OXR_OK;
}

// key_range of column `col` of batch `b`, measured on the first call and
// kept with the batch (knob agg.range_cache: 0 measures every time).
int key_range_of(oxr_stream *s, const oxr_batch *b, int col, KeyRange *out) {
  const bool cache = tune_get("agg.range_cache", 1) > 0;
  if (cache) {
    std::lock_guard<std::mutex> lock(b->stats_mu);
    for (const oxr_batch::KeyStats &k : b->key_stats)
      if (k.col == col) {
        out->known = true;
        out->min = k.min;
        out->count = k.count;
        return OXR_OK;
      }
  }
  OXR_TRY(key_range(s, b->cols[col], b->rows, out));
  if (cache) {
    std::lock_guard<std::mutex> lock(b->stats_mu);
    b->key_stats.push_back({col, out->min, out->count});
  }
  return OXR_OK;
}

// Grows the table so that `more` further distinct keys keep the load <= 1/2.
int ensure_capacity(oxr_stream *s, oxr_agg_state *st, uint64_t more) {
  uint64_t want = next_pow2(2 * (st->groups + more));
  if (st->mem && want <= st->t.cap) return OXR_OK;
  Table nt;
  void *nm = nullptr;
  OXR_TRY(table_alloc(s, st, std::max<uint64_t>(want, 16), &nt, &nm));
  if (st->mem) {
    const uint64_t slots = st->t.cap + 2;
    hipLaunchKernelGGL(rehash_kernel, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s->s, st->t,
                       nt, (int)st->aggs.size());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      stream_release(s, nm);
      return fail_hip(e, "rehash_kernel launch");
    }
    // the distinct count carries over
    e = hipMemcpyAsync(nt.distinct, st->t.distinct, 8, hipMemcpyDeviceToDevice, s->s);
    if (e == hipSuccess) e = hipMemcpyAsync(nt.seen, st->t.seen, 8, hipMemcpyDeviceToDevice, s->s);
    if (e != hipSuccess) {
      stream_release(s, nm);
      return fail_hip(e, "hipMemcpyAsync(rehash counters)");
    }
    stream_release(s, st->mem);
  }
  st->t = nt;
  st->mem = nm;
  return OXR_OK;
}

const char *func_name(int32_t f) {
  switch (f) {
    case OXR_SUM: return "SUM";
    case OXR_COUNT:
    case OXR_COUNT_STAR: return "COUNT";
    case OXR_MIN: return "MIN";
    case OXR_MAX: return "MAX";
  }
  return "?";
}

int compile_spec(const oxr_batch *like, const oxr_agg_spec *spec, oxr_agg_state *st) {
  if (spec->naggs < 0 || spec->naggs > kMaxAggs)
    return fail(OXR_EUNSUPPORTED, "rocm.aggregate",
                "at most " + std::to_string(kMaxAggs) + " aggregates per operator");
  if ((int)like->cols.size() > kMaxCols)
    return fail(OXR_EUNSUPPORTED, "rocm.aggregate", "too many input columns");
  st->key_col = spec->group_by;
  if (spec->group_by >= 0) {
    if (spec->group_by >= (int)like->cols.size())
      return fail(OXR_EPLAN, "rocm.aggregate", "group key column out of range");
    const int32_t kt = like->cols[spec->group_by].type;
    if (kt != OXR_INT64 && kt != OXR_INT32)
      return fail(OXR_EUNSUPPORTED, "cpu.int64_column",
                  "group key column has a type other than Int64/Int32");
    st->key_type = kt;
    st->key_name = like->names[spec->group_by];
    st->key_slot = st->prog.slot_of(spec->group_by);
  }
  OXR_TRY(compile_predicate(spec->filter, like->cols, &st->prog));
  if (spec->filter && spec->filter->count > 0 && spec->filter->nodes)
    st->filter_nodes.assign(spec->filter->nodes, spec->filter->nodes + spec->filter->count);
  for (int i = 0; i < spec->naggs; i++) {
    const oxr_agg &g = spec->aggs[i];
    st->funcs.push_back(g.func);
    if (g.func != OXR_COUNT_STAR && g.input.nodes && g.input.count > 0)
      st->input_nodes.emplace_back(g.input.nodes, g.input.nodes + g.input.count);
    else
      st->input_nodes.emplace_back();
  }
  for (int i = 0; i < spec->naggs; i++) {
    const oxr_agg &g = spec->aggs[i];
    AggDesc d{};
    std::string in_name = "*";
    int32_t in_type = OXR_INT64;
    if (g.func != OXR_COUNT_STAR) {
      ValueProg vp;
      OXR_TRY(compile_value(&g.input, like->cols, &st->prog, &vp));
      in_type = vp.type;
      d.vp = vp;
      d.nullable = vp.nullable;
      in_name = vp.col >= 0 ? like->names[vp.col] : "expr";
      if (in_type == OXR_F32_VECTOR)
        return fail(OXR_EUNSUPPORTED, "rocm.aggregate", "aggregate over a vector column");
    }
    d.in_type = (uint8_t)in_type;
    const bool fl = in_type == OXR_FLOAT32 || in_type == OXR_FLOAT64;
    int32_t out_t = in_type;
    switch (g.func) {
      case OXR_COUNT: d.kind = A_COUNT; out_t = OXR_INT64; break;
      case OXR_COUNT_STAR: d.kind = A_COUNT_STAR; out_t = OXR_INT64; break;
      case OXR_SUM:
        d.kind = fl ? A_SUM_F64 : A_SUM_I64;
        out_t = fl ? OXR_FLOAT64 : OXR_INT64;
        break;
      case OXR_MIN:
        d.kind = in_type == OXR_FLOAT64 ? A_MIN_F64 : in_type == OXR_FLOAT32 ? A_MIN_F32 : A_MIN_I64;
        break;
      case OXR_MAX:
        d.kind = in_type == OXR_FLOAT64 ? A_MAX_F64 : in_type == OXR_FLOAT32 ? A_MAX_F32 : A_MAX_I64;
        break;
      default:
        return fail(OXR_EPLAN, "rocm.aggregate", "unknown aggregate function");
    }
    st->aggs.push_back(d);
    st->out_types.push_back(out_t);
    st->names.push_back(g.name ? std::string(g.name)
                               : std::string(func_name(g.func)) + "(" + in_name + ")");
  }
  for (auto &c : like->cols) st->col_types.push_back(c.type);
  st->hint = spec->groups_hint;
  return OXR_OK;
}

// The group of every slot of a unique-key join table: Int32 groups as
// (state << 32) | group, Int64 groups as {group, state}; state 0 = no build
// row, 1 = a group, 2 = the NULL group.
__global__ void gos_kernel(JTable t, DCol grp, uint64_t nslots, uint64_t *gos32, uint64_t *gos64) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  uint64_t state = 0;
  int64_t g = 0;
  if (t.count[i]) {
    const uint64_t b = t.rows[t.start[i]];
    const bool v = !grp.valid || ((grp.valid[b >> 5] >> (b & 31)) & 1u);
    g = grp.type == OXR_INT32 ? (int64_t)((const int32_t *)grp.values)[b] : ((const int64_t *)grp.values)[b];
    state = v ? 1 : 2;
    if (!v) g = 0;
  }
  if (gos32) gos32[i] = (state << 32) | (uint32_t)g;
  else {
    gos64[2 * i] = (uint64_t)g;
    gos64[2 * i + 1] = state;
  }
}

// ------------------------------------------------------------------- spill
// Rows the generated kernel's block tables do not hold (the cold tail of a
// high-cardinality key) are spilled as records (the group index, then the
// row's contributions in as few 32-bit words as their types allow), each
// block appending to 32 coarse
// buckets (their tails stay in L2 while they fill, so the writes reach HBM
// as whole lines), with a histogram over partitions of 2^slog2 groups.
// Random HBM atomics run at ~24 G/s on this device whatever their scope or
// footprint, so instead of one atomic per row and aggregate the records are
// partitioned (scatter: one coarse bucket of one block per workgroup, a few
// fine partitions each) and each partition is folded in LDS by one
// workgroup (fold), which adds one update per (group, aggregate) to the HBM
// table. The records' layout depends on the query's aggregates (SpillRec),
// so both passes are generated with the aggregation kernel (spill_scatter,
// spill_fold).

struct FoldKinds {
  uint8_t kind[kMaxAggs];
};

// ------------------------------------------------------------- hot groups
// Where a sampled row's group comes from: the key column itself, or the
// probe key through a unique-key join (the group of each join table slot in
// gos32 / gos64, as the generated kernel reads it).
struct HotKeySrc {
  DCol key;
  int32_t join;
  JTable jt;
  const uint64_t *gos32;
  const uint64_t *gos64;
};

// The group index (key - smin) of every `stride`-th row into grp[i], or
// 0xFFFFFFFF when the row has no group in [smin, smin + srange): the sample
// the hot map is picked from. (Counting with atomics instead serialises on
// the hottest group: a Zipf key's top group is a tenth of the sample.)
__global__ void hot_sample_kernel(HotKeySrc src, uint64_t samples, uint64_t stride, int64_t smin,
                                  uint64_t srange, uint32_t *grp) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= samples) return;
  grp[i] = 0xFFFFFFFFu;
  const uint64_t r = i * stride;  // host: (samples - 1) * stride < rows
  const DCol &c = src.key;
  if (c.valid && !((c.valid[r >> 5] >> (r & 31)) & 1u)) return;
  int64_t k = c.type == OXR_INT32 ? (int64_t)((const int32_t *)c.values)[r] : ((const int64_t *)c.values)[r];
  if (src.join) {
    const uint64_t slot = join_find(src.jt, k);
    if (slot == kNoSlot) return;
    uint32_t state;
    if (src.gos32) {
      const uint64_t e = src.gos32[slot];
      state = (uint32_t)(e >> 32);
      k = (int64_t)(int32_t)(uint32_t)e;
    } else {
      k = (int64_t)src.gos64[2 * slot];
      state = (uint32_t)src.gos64[2 * slot + 1];
    }
    if ((state & 0xFFu) != 1) return;  // no build row, or its group is NULL
  }
  if (k < smin || (uint64_t)k - (uint64_t)smin >= srange) return;
  grp[i] = (uint32_t)((uint64_t)k - (uint64_t)smin);
}

// Stamps each join table slot's group entry with its group's hot id (the
// state word's bits 8 and up: hot id + 1, 0 for a cold group), so the hot
// kernel's one lookup gives group and hot id together.
__global__ void gos_hot_kernel(uint64_t *gos32, uint64_t *gos64, uint64_t nslots, const int32_t *hotmap, int64_t smin,
                               uint64_t srange) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= nslots) return;
  int64_t g;
  uint32_t state;
  if (gos32) {
    g = (int64_t)(int32_t)(uint32_t)gos32[i];
    state = (uint32_t)(gos32[i] >> 32);
  } else {
    g = (int64_t)gos64[2 * i];
    state = (uint32_t)gos64[2 * i + 1];
  }
  if (state != 1 || g < smin || (uint64_t)g - (uint64_t)smin >= srange) return;
  const uint32_t h1 = (uint32_t)(hotmap[(uint64_t)g - (uint64_t)smin] + 1);
  if (!h1) return;
  if (gos32) gos32[i] |= (uint64_t)h1 << 40;
  else gos64[2 * i + 1] = state | (h1 << 8);
}

__global__ void fill_i32_kernel(int32_t *v, uint64_t n, int32_t x) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = x;
}

// Hot id h -> the h-th most sampled group (sorted descending), for groups
// sampled at least `min_count` times; hotkeys[h] = its key (smin + index).
// Groups past srange (no group) and runs past the run count (zero counts)
// get no hot id.
__global__ void hot_assign_kernel(const uint32_t *cnt_sorted, const uint32_t *grp_sorted, uint32_t nhot,
                                  uint32_t min_count, int64_t smin, uint64_t srange, int32_t *hotmap,
                                  int64_t *hotkeys) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nhot) return;
  const uint32_t g = grp_sorted[h];
  hotkeys[h] = (int64_t)((uint64_t)smin + g);
  if (cnt_sorted[h] >= min_count && g < srange) hotmap[g] = (int32_t)h;
}

// Folds the blocks' hot partials into the HBM table: block (x, y) takes 64
// hot ids and the source blocks y, y + gridDim.y, ...; 4 threads per hot id
// split those, combine in LDS, and one update per (hot id, aggregate) that
// any of its source blocks used goes to the table.
__global__ __launch_bounds__(256) void hot_reduce_kernel(const int64_t *part, const uint8_t *flags, uint32_t nblocks,
                                                         uint32_t nhot, const int64_t *hotkeys, Table t, FoldKinds fk,
                                                         int na) {
  __shared__ int64_t s_v[4][kMaxAggs][64];
  __shared__ uint32_t s_f[4][64];
  const int hl = threadIdx.x & 63, q = threadIdx.x >> 6;
  const uint32_t h = blockIdx.x * 64 + hl;
  int64_t v[kMaxAggs];
  for (int a = 0; a < na; a++) v[a] = acc_init(fk.kind[a]);
  uint32_t f = 0;
  if (h < nhot) {
    for (uint32_t b = blockIdx.y + gridDim.y * q; b < nblocks; b += 4 * gridDim.y) {
      const uint32_t fb = flags[(uint64_t)b * nhot + h];
      if (!(fb & 1u)) continue;
      f |= fb;
      for (int a = 0; a < na; a++) v[a] = combine_kind(fk.kind[a], v[a], part[((uint64_t)b * na + a) * nhot + h]);
    }
  }
  for (int a = 0; a < na; a++) s_v[q][a][hl] = v[a];
  s_f[q][hl] = f;
  __syncthreads();
  if (q != 0 || h >= nhot) return;
  for (int k = 1; k < 4; k++) {
    f |= s_f[k][hl];
    for (int a = 0; a < na; a++) v[a] = combine_kind(fk.kind[a], v[a], s_v[k][a][hl]);
  }
  if (!(f & 1u)) return;
  bool ok = true;
  const uint64_t gs = global_slot(t, hotkeys[h], &ok);
  if (!ok) {
    atomicOr(t.overflow, 1u);
    return;
  }
  for (int a = 0; a < na; a++) {
    if (v[a] != acc_init(fk.kind[a]) || fk.kind[a] == A_SUM_F64) acc_apply_global(fk.kind[a], &t.acc[a][gs], v[a]);
    if (t.has[a] && ((f >> (1 + a)) & 1u)) t.has[a][gs] = 1;
  }
}

// The join an update folds through (oxr_agg_update_join), or none.
struct JoinSide {
  const oxr_join_table *table = nullptr;
  const oxr_batch *build = nullptr;
  int32_t build_group = -1;
  int32_t probe_slot = -1;
  int32_t probe_key = -1;
};

// The generator's view of an aggregation state: with probe_key >= 0 the key
// comes through a join on that probe column, with -2 from the state's own key.
struct ShapeOf {
  AggJitShape shape;
  oxr_expr fexpr;
  std::vector<oxr_expr> inputs;
  ShapeOf(const oxr_agg_state *st, int32_t probe_key) : inputs(st->aggs.size()) {
    fexpr = oxr_expr{st->filter_nodes.data(), (int32_t)st->filter_nodes.size()};
    shape.filter = st->filter_nodes.empty() ? nullptr : &fexpr;
    shape.join = probe_key >= 0;
    shape.key_col = shape.join ? probe_key : st->key_col;
    for (size_t a = 0; a < st->aggs.size(); a++) {
      shape.kinds.push_back(st->aggs[a].kind);
      shape.nullable.push_back(st->aggs[a].nullable);
      inputs[a] = oxr_expr{st->input_nodes[a].data(), (int32_t)st->input_nodes[a].size()};
      shape.inputs.push_back(st->aggs[a].kind == A_COUNT_STAR ? nullptr : &inputs[a]);
    }
  }
};

// Samples the update's input (every stride-th row, about 1M rows), counts
// each sampled group of the spill range, and gives the most sampled ones
// (up to `cap`, each sampled at least twice) dense hot ids: A.hotmap,
// A.nhot, and the per-block partials A.hot_part / A.hot_flags the hot
// kernel writes. *mem holds it all (released by the caller after the
// reduce); *nhot = 0 when there is nothing to do.
int build_hot_map(oxr_stream *s, const oxr_agg_state *st, const oxr_batch *in, const JoinSide *js,
                  const KeyRange *kr, AggJitArgs &A, uint32_t grid, uint32_t cap, void **mem, int64_t **hotkeys,
                  uint32_t *nhot) {
  const uint64_t srange = kr->count;
  const int na = (int)st->aggs.size();
  const uint64_t n = in->rows;
  const uint64_t samples = std::min<uint64_t>(n, 1ull << 20);
  const uint32_t nh = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(cap, srange), samples);
  const uint64_t stride = n / samples;  // (samples - 1) * stride < n
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  // The map depends only on the batch's key column (through the join), the
  // key range and its size, not on the aggregation's filter or inputs: it is
  // kept with the batch and every later update over it takes it (knob
  // agg.hot_cache). Each update still gets its own partials and flags.
  const bool cache = tune_get("agg.hot_cache", 1) > 0;
  const uint64_t jser = js ? js->table->serial : 0;
  const int32_t kcol = js ? js->probe_key : st->key_col, bgrp = js ? js->build_group : -1;
  const size_t b_part = up((size_t)grid * na * nh * 8), b_flags = up((size_t)grid * nh);
  if (cache) {
    std::lock_guard<std::mutex> lock(in->stats_mu);
    for (const oxr_batch::HotMap &h : in->hot_maps)
      if (h.join_serial == jser && h.key_col == kcol && h.build_group == bgrp && h.smin == kr->min &&
          h.srange == srange && h.nh == nh) {
        // built on another stream, perhaps: its kernels come first
        OXR_HIP(hipStreamWaitEvent(s->s, h.ready, 0));
        OXR_TRY(stream_alloc(s, b_part + b_flags, mem));
        A.hot_part = (int64_t *)*mem;
        A.hot_flags = (uint8_t *)((char *)*mem + b_part);
        A.hotmap = h.hotmap;
        A.nhot = nh;
        *hotkeys = h.hotkeys;
        *nhot = nh;
        return OXR_OK;
      }
  }
  // the sampled groups, sorted, run-length encoded into (group, count)
  // runs, the runs sorted by count (descending): no atomics
  size_t b_sort1 = 0, b_rle = 0, b_sort2 = 0;
  const int ns = (int)samples;
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(nullptr, b_sort1, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                   ns, 0, 32, s->s);
  if (e == hipSuccess)
    e = hipcub::DeviceRunLengthEncode::Encode(nullptr, b_rle, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr, ns, s->s);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b_sort2, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                     (const uint32_t *)nullptr, (uint32_t *)nullptr, ns, 0, 32,
                                                     s->s);
  if (e != hipSuccess) return fail_hip(e, "hot map sizes");
  const size_t b_s = up(samples * 4);
  const size_t b_tmp = up(std::max(b_sort1, std::max(b_rle, b_sort2)));
  const size_t b_map = up(srange * 4) + up((size_t)nh * 8);  // the map and the hot keys
  // the map in an allocation of its own when it is kept with the batch
  void *pmem = nullptr;
  if (cache) OXR_TRY(stream_alloc(s, b_map, &pmem));
  const size_t total = 6 * b_s + 256 + (cache ? 0 : b_map) + b_part + b_flags + b_tmp;
  int arc = stream_alloc(s, total, mem);
  if (arc != OXR_OK) {
    if (pmem) stream_release(s, pmem);
    return arc;
  }
  char *q = (char *)*mem;
  uint32_t *grp = (uint32_t *)q, *grp_sorted = (uint32_t *)(q + b_s), *runs = (uint32_t *)(q + 2 * b_s),
           *counts = (uint32_t *)(q + 3 * b_s), *cnt_sorted = (uint32_t *)(q + 4 * b_s),
           *run_sorted = (uint32_t *)(q + 5 * b_s), *nruns = (uint32_t *)(q + 6 * b_s);
  q += 6 * b_s + 256;
  char *mq = cache ? (char *)pmem : q;
  if (!cache) q += b_map;
  int32_t *hotmap = (int32_t *)mq;
  *hotkeys = (int64_t *)(mq + up(srange * 4));
  A.hot_part = (int64_t *)q;
  q += b_part;
  A.hot_flags = (uint8_t *)q;
  q += b_flags;
  void *tmp = q;
  // counts past the run count stay 0 (sorted last, never hot)
  int rc = fill_zero(s, counts, samples * 4);
  if (rc != OXR_OK) {
    if (pmem) stream_release(s, pmem);
    return rc;
  }
  const unsigned tb = 256;
  ProfScope prof(s->s, "aggregate_hot_map");  // a map built (not one taken from the batch)
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)std::min<uint64_t>((srange + tb - 1) / tb, 4096)), dim3(tb), 0,
                     s->s, hotmap, srange, -1);
  HotKeySrc src{};
  if (js) {
    const DevCol &k = in->cols[js->probe_key];
    src.key.values = k.values;
    src.key.valid = (const uint32_t *)k.validity;
    src.key.type = k.type;
    src.join = 1;
    src.jt = A.jt;
    src.gos32 = A.gos;
    src.gos64 = A.gos64;
  } else {
    const DevCol &k = in->cols[st->key_col];
    src.key.values = k.values;
    src.key.valid = (const uint32_t *)k.validity;
    src.key.type = k.type;
  }
  hipLaunchKernelGGL(hot_sample_kernel, dim3((unsigned)((samples + tb - 1) / tb)), dim3(tb), 0, s->s, src, samples,
                     stride, kr->min, srange, grp);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(tmp, b_sort1, grp, grp_sorted, ns, 0, 32, s->s);
  if (e == hipSuccess)
    e = hipcub::DeviceRunLengthEncode::Encode(tmp, b_rle, grp_sorted, runs, counts, nruns, ns, s->s);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairsDescending(tmp, b_sort2, counts, cnt_sorted, runs, run_sorted, ns, 0, 32,
                                                     s->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(hot_assign_kernel, dim3((nh + tb - 1) / tb), dim3(tb), 0, s->s, cnt_sorted, run_sorted, nh,
                       2u, kr->min, srange, hotmap, *hotkeys);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    if (pmem) stream_release(s, pmem);
    return fail_hip(e, "hot map");
  }
  A.hotmap = hotmap;
  A.nhot = nh;
  *nhot = nh;
  if (cache) {
    hipEvent_t ready = nullptr;
    e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ready, s->s);
    if (e != hipSuccess) {
      if (ready) (void)hipEventDestroy(ready);
      // not kept: this update's map is released with its scratch
      (void)hipGetLastError();
      stream_release(s, pmem);
      return fail_hip(e, "hot map event");
    }
    std::lock_guard<std::mutex> lock(in->stats_mu);
    oxr_batch::HotMap h;
    h.ready = ready;
    h.join_serial = jser;
    h.key_col = kcol;
    h.build_group = bgrp;
    h.smin = kr->min;
    h.srange = srange;
    h.nh = nh;
    h.mem = pmem;
    h.hotmap = hotmap;
    h.hotkeys = *hotkeys;
    h.core = retain(s->core);
    in->hot_maps.push_back(h);
  }
  return OXR_OK;
}

// The update as the kernel generated for this aggregation's shape, when the
// generator covers it (a plain key or one global group, or a join on unique
// build keys). OXR_EUNSUPPORTED: the interpreter kernel runs it.
int run_update_generated(oxr_stream *s, oxr_agg_state *st, const oxr_batch *in, const JoinSide *js,
                         uint32_t *err, const KeyRange *kr) {
  if (tune_get("agg.jit", 1) == 0 || st->aggs.empty() || (js && !js->table->t.unique))
    return OXR_EUNSUPPORTED;
  ShapeOf so(st, js ? js->probe_key : -2);
  JitAgg ja;
  const AggJitShape &sh = so.shape;
  const int jrc = jit_aggregate(s->dev, in->cols, sh, &ja);
  if (jrc != OXR_OK) return jrc;
  AggJitArgs A;
  std::memset(&A, 0, sizeof A);
  A.n = in->rows;
  for (size_t i = 0; i < in->cols.size(); i++) {
    A.col[i] = in->cols[i].values;
    A.valid[i] = (const uint32_t *)in->cols[i].validity;
  }
  A.t = st->t;
  A.error = err;
  std::copy(ja.lits.begin(), ja.lits.end(), A.lit);
  void *gos = nullptr;
  if (js) {
    // the group of every slot of the join table
    const JTable &t = js->table->t;
    const DevCol &g = js->build->cols[js->build_group];
    const uint64_t nslots = t.slots + 1;
    const bool g32 = g.type == OXR_INT32;
    OXR_TRY(stream_alloc(s, nslots * (g32 ? 8 : 16), &gos));
    DCol gc;
    gc.values = g.values;
    gc.valid = (const uint32_t *)g.validity;
    gc.type = g.type;
    gc.dim = 0;
    hipLaunchKernelGGL(gos_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s->s, t, gc, nslots,
                       g32 ? (uint64_t *)gos : nullptr, g32 ? nullptr : (uint64_t *)gos);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      stream_release(s, gos);
      return fail_hip(e, "gos_kernel launch");
    }
    A.jt = t;
    A.gos = g32 ? (const uint64_t *)gos : nullptr;
    A.gos64 = g32 ? nullptr : (const uint64_t *)gos;
  }
  // a small measured key range: the dense kernel (a per-lane LDS column of
  // accumulators for every (key, aggregate) pair, its lo parts in exact mode)
  hipFunction_t fn = ja.fn;
  if (ja.dense && !js && kr && kr->known && kr->count > 0 && kr->count <= 32 && tune_get("agg.dense", 1) > 0) {
    const uint64_t nkeys = kr->count + (in->cols[st->key_col].validity ? 1 : 0);
    const uint64_t need = nkeys * st->aggs.size() * (tune_get("agg.exact_f64", 0) > 0 ? 2 : 1);
    if (nkeys * st->aggs.size() <= 32 && need <= 32) {
      fn = need <= 16 ? ja.dense16 : ja.dense32;
      A.dense_min = kr->min;
      A.dense_keys = (uint32_t)kr->count;
      A.dense_null = in->cols[st->key_col].validity ? 1 : 0;
    }
  }
  // a global aggregate: the keyless kernel (knob agg.keyless)
  if (ja.keyless && !js && tune_get("agg.keyless", 1) > 0) fn = ja.keyless_fn;
  // spill the cold tail of a high-cardinality key (see spill_fold):
  // a measured key range that the fold's partitions cover, every aggregate
  // in the block table, plain FP64 sums, and enough rows to pay for it
  const int na = (int)st->aggs.size();
  const bool exact = tune_get("agg.exact_f64", 0) > 0;
  uint32_t slog2 = 0, nparts = 0;
  int nnull = 0;
  for (int a = 0; a < na; a++) nnull += st->aggs[a].nullable != 0;
  if (fn == ja.fn && ja.spill && kr && kr->known && kr->count > 0 && !exact && na <= kLdsAggs &&
      tune_get("agg.spill", 1) > 0 && in->rows >= (uint64_t)std::max<int64_t>(0, tune_get("agg.spill_rows", 1 << 20))) {
    slog2 = fold_log2(na, nnull);
    const uint64_t parts = (kr->count + (1ull << slog2) - 1) >> slog2;
    if (parts <= (uint64_t)kMaxSpillParts && kr->count < (1ull << 32)) nparts = (uint32_t)parts;
  }
  // hot groups: with the spill on (a wide key range) over enough rows, a
  // sample of the input picks the groups that get dense LDS accumulators
  // (aggregate_hot) instead of the hashed block table
  bool any_nullable = false;
  for (int a = 0; a < na; a++) any_nullable |= st->aggs[a].nullable != 0;
  const uint32_t hot_cap = (uint32_t)hot_capacity(na, any_nullable);
  const bool hot_try = nparts && ja.hot && fn == ja.fn && tune_get("agg.hot", 1) > 0 && kr->count > 4ull * hot_cap &&
                       in->rows >= (uint64_t)std::max<int64_t>(1, tune_get("agg.hot_rows", 1 << 22));
  // workgroups per CU: the hot kernel hides its LDS atomics' latency with
  // more resident blocks (knob agg.hot_grid), the others take agg.jit_grid
  const uint64_t tile = 256ull * ja.rows_per_lane;
  const uint64_t ntiles = (in->rows + tile - 1) / tile;
  const uint64_t per_cu = (uint64_t)std::max<int64_t>(
      1, hot_try ? tune_get("agg.hot_grid", 8)
                 : fn == ja.keyless_fn && fn ? tune_get("agg.keyless_grid", 4) : tune_get("agg.jit_grid", 4));
  const uint64_t grid = std::min<uint64_t>(ntiles, (uint64_t)std::max(1, s->dev->cu_count) * per_cu);
  const uint32_t words = (uint32_t)ja.spill_words;  // 32-bit words per record
  void *spill_mem = nullptr;
  uint32_t *spill = nullptr;
  uint32_t *spill_n = nullptr, *hist = nullptr, *offs = nullptr;
  uint64_t cap = 0;
  const uint32_t sper = nparts ? (nparts + kSpillBuckets - 1) / kSpillBuckets : 1;
  if (nparts) {
    // a block's 32 buckets together hold a fraction (knob agg.spill_pct,
    // percent) of its rows; rows past a full bucket take the HBM path
    const uint64_t tiles_per_block = (ntiles + grid - 1) / grid;
    const int64_t pct = std::min<int64_t>(400, std::max<int64_t>(1, tune_get("agg.spill_pct", 60)));
    cap = std::max<uint64_t>(64, (tiles_per_block * tile * (uint64_t)pct / 100 + kSpillBuckets - 1) / kSpillBuckets);
    // the kernel addresses a block's buckets with 32-bit offsets
    cap = std::min<uint64_t>(cap, (uint64_t)(UINT32_MAX / 2) / (kSpillBuckets * words));
    const uint64_t recs = grid * kSpillBuckets * cap;
    const uint64_t nh = (uint64_t)nparts * grid + 1;
    const size_t bytes = recs * words * 4 + (grid * kSpillBuckets + 2 * nh) * 4 + 1024;
    if (stream_alloc(s, bytes, &spill_mem) == OXR_OK) {
      char *q = (char *)spill_mem;
      spill = (uint32_t *)q;
      q += (recs * words * 4 + 255) / 256 * 256;
      spill_n = (uint32_t *)q;
      q += grid * kSpillBuckets * 4;
      hist = (uint32_t *)q;
      q += nh * 4;
      offs = (uint32_t *)q;
      const int zrc = fill_zero(s, hist + nh - 1, 4);  // the scan's trailing zero: offs[nh - 1] = total
      if (zrc != OXR_OK) {
        stream_release(s, spill_mem);
        stream_release(s, gos);
        return zrc;
      }
      A.spill = spill;
      A.spill_cap = cap;
      A.spill_n = spill_n;
      A.hist = hist;
      A.smin = kr->min;
      A.srange = kr->count;
      A.slog2 = slog2;
      A.sparts = nparts;
      A.sper = sper;
    } else {
      clear_error();  // no room for the spill buffers: the HBM path takes every row
      nparts = 0;
    }
  }
  void *hot_mem = nullptr;
  int64_t *hotkeys = nullptr;
  uint32_t nhot = 0;
  int rc = OXR_OK;
  if (nparts && hot_try) {
    rc = build_hot_map(s, st, in, js, kr, A, (uint32_t)grid, hot_cap, &hot_mem, &hotkeys, &nhot);
    if (rc == OXR_OK && nhot) fn = ja.hot_fn;
    if (rc == OXR_OK && nhot && js) {
      // the join's group entries carry the hot ids: one lookup per row
      const uint64_t nslots = js->table->t.slots + 1;
      hipLaunchKernelGGL(gos_hot_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, s->s,
                         A.gos ? (uint64_t *)gos : nullptr, A.gos64 ? (uint64_t *)gos : nullptr, nslots, A.hotmap,
                         kr->min, kr->count);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail_hip(e, "gos_hot_kernel launch");
      A.gos_hot = 1;
    }
  }
  if (rc == OXR_OK) {
    ProfScope prof(s->s, fn == ja.fn ? "aggregate"
                         : fn == ja.hot_fn ? "aggregate_hot"
                         : fn == ja.keyless_fn ? "aggregate_keyless"
                                               : "aggregate_dense");
    rc = jit_launch_raw(fn, (unsigned)grid, s->s, &A, sizeof A);
  }
  if (rc == OXR_OK && nhot) {
    ProfScope prof(s->s, "aggregate_hot_reduce");
    FoldKinds fk{};
    for (int a = 0; a < na; a++) fk.kind[a] = st->aggs[a].kind;
    const dim3 g((unsigned)((nhot + 63) / 64), 8), b(256);
    hipLaunchKernelGGL(hot_reduce_kernel, g, b, 0, s->s, A.hot_part, A.hot_flags, (uint32_t)grid, nhot, hotkeys,
                       st->t, fk, na);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "hot_reduce_kernel launch");
  }
  if (hot_mem) stream_release(s, hot_mem);
  if (rc == OXR_OK && nparts) {
    ProfScope prof(s->s, "aggregate_spill");
    const uint64_t nh = (uint64_t)nparts * grid + 1;
    size_t tmp_bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, hist, offs, (int)nh, s->s);
    void *tmp = nullptr;
    if (e == hipSuccess) rc = stream_alloc(s, tmp_bytes + 16, &tmp);
    if (rc == OXR_OK && e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, hist, offs, (int)nh, s->s);
    // the partitioned records, sized by the spill's total
    uint32_t total = 0;
    uint32_t *pin_total = nullptr;
    if (rc == OXR_OK && e == hipSuccess) rc = stream_pinned(s, 4, (void **)&pin_total);
    if (rc == OXR_OK && e == hipSuccess) e = hipMemcpyAsync(pin_total, offs + nh - 1, 4, hipMemcpyDeviceToHost, s->s);
    if (rc == OXR_OK && e == hipSuccess) e = hipStreamSynchronize(s->s);
    if (rc == OXR_OK && e == hipSuccess) total = *pin_total;
    void *sorted_mem = nullptr;
    if (rc == OXR_OK && e == hipSuccess) rc = stream_alloc(s, (size_t)std::max<uint32_t>(total, 1) * words * 4, &sorted_mem);
    SpillArgs sa;
    std::memset(&sa, 0, sizeof sa);
    sa.spill = spill;
    sa.spill_n = spill_n;
    sa.offs = offs;
    sa.recs = (uint32_t *)sorted_mem;
    sa.cap = cap;
    sa.B = (uint32_t)grid;
    sa.nparts = nparts;
    sa.sper = sper;
    sa.slog2 = slog2;
    sa.smin = kr->min;
    sa.t = st->t;
    sa.error = err;
    if (rc == OXR_OK && e == hipSuccess) {
      ProfScope p2(s->s, "aggregate_scatter");
      rc = jit_launch_raw(ja.scatter_fn, (unsigned)(grid * kSpillBuckets), s->s, &sa, sizeof sa, 256);
    }
    if (rc == OXR_OK && e == hipSuccess) {
      ProfScope p3(s->s, "aggregate_fold");
      rc = jit_launch_raw(ja.fold_fn, nparts, s->s, &sa, sizeof sa, (unsigned)kFoldThreads);
    }
    if (rc == OXR_OK && e != hipSuccess) rc = fail_hip(e, "aggregate spill passes");
    stream_release(s, tmp);
    stream_release(s, sorted_mem);
  }
  if (spill_mem) stream_release(s, spill_mem);
  stream_release(s, gos);
  return rc;
}

int poisoned_error() {
  return fail(OXR_EEXECUTION, "rocm.aggregate", "an earlier update of this aggregation failed part way");
}

int run_update(oxr_stream *s, oxr_agg_state *st, const oxr_batch *in, const JoinSide *js = nullptr) {
  if (st->poisoned) return poisoned_error();
  if (in->cols.size() != st->col_types.size())
    return fail(OXR_EPLAN, "rocm.aggregate", "batch does not match the aggregation's schema");
  for (size_t i = 0; i < in->cols.size(); i++)
    if (in->cols[i].type != st->col_types[i])
      return fail(OXR_EPLAN, "rocm.aggregate", "batch column types differ from the spec's");
  const uint64_t n = in->rows;
  if (n == 0) return OXR_OK;
  // keep the load at most one half for every key this batch might add: at
  // most a key per row, per build row of a join, and per value of the key's
  // range (one pass over the key column measures it)
  uint64_t more = st->key_col == -1 ? 1 : n;
  if (js) more = std::min<uint64_t>(more, std::max<uint64_t>(js->table->nrows, 1));
  // (groups_hint only pre-sizes the table at create: capping `more` with it
  // could overflow the table part way through an update)
  KeyRange kr;
  if (st->key_col != -1 && n > (uint64_t)std::max<int64_t>(0, tune_get("agg.range_rows", kRangeCheckRows))) {
    if (js) OXR_TRY(key_range_of(s, js->build, js->build_group, &kr));
    else OXR_TRY(key_range_of(s, in, st->key_col, &kr));
    more = std::min<uint64_t>(more, std::max<uint64_t>(kr.count, 1));
  }
  OXR_TRY(ensure_capacity(s, st, more));

  std::unique_ptr<Params> pp(new Params());
  Params &P = *pp;
  std::memset(&P, 0, sizeof P);
  P.n = n;
  for (size_t i = 0; i < in->cols.size(); i++) {
    const DevCol &c = in->cols[i];
    if (n && ((uintptr_t)c.values % 16) != 0)
      return fail(OXR_EUNSUPPORTED, "rocm.aggregate", "column values must be 16-byte aligned");
    P.cols[i].values = c.values;
    P.cols[i].valid = (const uint32_t *)c.validity;
    P.cols[i].type = c.type;
    P.cols[i].dim = c.dim;
  }
  std::copy(st->prog.ins.begin(), st->prog.ins.end(), P.ins);
  std::copy(st->prog.leaves.begin(), st->prog.leaves.end(), P.leaves);
  std::copy(st->prog.bprog.begin(), st->prog.bprog.end(), P.bprog);
  std::copy(st->prog.magic.begin(), st->prog.magic.end(), P.magic);
  for (size_t i = 0; i < st->prog.slot_cols.size(); i++) P.slot_cols[i] = (int8_t)st->prog.slot_cols[i];
  P.nslots = (int32_t)st->prog.slot_cols.size();
  P.nb = (int32_t)st->prog.bprog.size();
  P.key_slot = st->key_col >= 0 ? st->key_slot : -1;
  P.naggs = (int32_t)st->aggs.size();
  P.lds_aggs = P.naggs <= kLdsAggs ? P.naggs : 0;
  P.exact_f64 = tune_get("agg.exact_f64", 0) > 0 ? 1 : 0;
  for (int a = 0; a < P.naggs; a++) P.aggs[a] = st->aggs[a];
  P.t = st->t;
  if (js) {
    const DevCol &g = js->build->cols[js->build_group];
    P.join = 1;
    P.probe_slot = js->probe_slot;
    P.jt = js->table->t;
    P.build_group.values = g.values;
    P.build_group.valid = (const uint32_t *)g.validity;
    P.build_group.type = g.type;
  }
  void *err = nullptr;
  OXR_TRY(stream_alloc(s, 64, &err));
  P.error = (uint32_t *)err;
  int rc = fill_zero(s, err, 64);
  const uint64_t ntiles = (n + kTileRows - 1) / kTileRows;
  const uint64_t per_cu = (uint64_t)std::max<int64_t>(1, tune_get("agg.grid", 8));
  const uint64_t grid = std::min<uint64_t>(ntiles, (uint64_t)std::max(1, s->dev->cu_count) * per_cu);
  bool generated = false;
  // from here on a failure may leave some of the batch's rows applied
  st->poisoned = true;
  if (rc == OXR_OK) {
    const int grc = run_update_generated(s, st, in, js, (uint32_t *)err, &kr);
    if (grc == OXR_OK) generated = true;
    else if (grc != OXR_EUNSUPPORTED) rc = grc;
  }
  if (rc == OXR_OK && !generated) {
    ProfScope prof(s->s, "aggregate_vm");
    switch (slot_bucket(P.nslots)) {
      case 2: hipLaunchKernelGGL(aggregate_kernel<2>, dim3((unsigned)grid), dim3(kBlock), 0, s->s, P); break;
      case 4: hipLaunchKernelGGL(aggregate_kernel<4>, dim3((unsigned)grid), dim3(kBlock), 0, s->s, P); break;
      default: hipLaunchKernelGGL(aggregate_kernel<8>, dim3((unsigned)grid), dim3(kBlock), 0, s->s, P); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "aggregate_kernel launch");
  }
  uint64_t host[4] = {0, 0, 0, 0};  // errors (divide by zero | spill fold << 32), distinct, overflow
  uint64_t *pin = nullptr;
  if (rc == OXR_OK) rc = stream_pinned(s, 32, (void **)&pin);
  if (rc == OXR_OK) {
    pin[0] = pin[1] = pin[2] = 0;
    hipError_t e = hipMemcpyAsync(&pin[0], err, 8, hipMemcpyDeviceToHost, s->s);
    if (e == hipSuccess) e = hipMemcpyAsync(&pin[1], st->t.distinct, 8, hipMemcpyDeviceToHost, s->s);
    if (e == hipSuccess) e = hipMemcpyAsync(&pin[2], st->t.overflow, 4, hipMemcpyDeviceToHost, s->s);
    if (e == hipSuccess) e = hipStreamSynchronize(s->s);
    if (e != hipSuccess) rc = fail_hip(e, "aggregate result");
    else std::memcpy(host, pin, 24);
  }
  stream_release(s, err);
  if (rc != OXR_OK) return rc;
  if ((uint32_t)host[0]) return fail(OXR_EEXECUTION, "rocm.aggregate", "Divide by zero error");
  if ((uint32_t)(host[0] >> 32))
    return fail(OXR_EEXECUTION, "rocm.aggregate", "internal: the spill fold's partitions exceed its LDS table");
  // cannot happen while the sizing above holds (load <= 1/2 for every key
  // the batch can add); never a fallback signal: rows have been applied
  if ((uint32_t)host[2]) return fail(OXR_EEXECUTION, "rocm.aggregate", "the group table overflowed");
  st->groups = host[1];
  st->poisoned = false;
  return OXR_OK;
}

int run_finish(oxr_stream *s, oxr_agg_state *st, oxr_batch **out) {
  if (st->poisoned) return poisoned_error();
  if (!st->mem) OXR_TRY(ensure_capacity(s, st, 1));
  const Table &t = st->t;
  // ordinary keys
  const uint64_t cap = t.cap;
  const uint64_t g = st->groups;
  // the key sort takes a 32-bit signed item count
  if (g >= (uint64_t)INT32_MAX) return fail(OXR_EUNSUPPORTED, "rocm.aggregate", "2^31 - 1 groups or more");
  void *buf = nullptr;
  const size_t gb = (g ? g : 1) * 8;
  OXR_TRY(stream_alloc(s, gb * 4 + 64, &buf));
  int64_t *k_in = (int64_t *)buf;
  uint64_t *s_in = (uint64_t *)((char *)buf + gb);
  int64_t *k_out = (int64_t *)((char *)buf + 2 * gb);
  uint64_t *s_out = (uint64_t *)((char *)buf + 3 * gb);
  uint64_t *count = (uint64_t *)((char *)buf + 4 * gb);
  int rc = fill_zero(s, count, 8);
  if (rc == OXR_OK && g) {
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s->s, t, k_in,
                       s_in, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "compact_kernel launch");
  }
  uint32_t seen[2] = {0, 0};
  uint32_t *pin_seen = nullptr;  // read after the wait below
  if (rc == OXR_OK) rc = stream_pinned(s, 8, (void **)&pin_seen);
  if (rc == OXR_OK) {
    hipError_t e = hipMemcpyAsync(pin_seen, t.seen, 8, hipMemcpyDeviceToHost, s->s);
    if (e != hipSuccess) rc = fail_hip(e, "aggregate seen flags");
  }
  // sort ordinary keys ascending (stable radix sort of signed keys)
  void *tmp = nullptr;
  if (rc == OXR_OK && g > 1) {
    size_t tmp_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k_in, k_out, s_in, s_out,
                                                      (int)g, 0, 64, s->s);
    if (e == hipSuccess) {
      rc = stream_alloc(s, tmp_bytes + 16, &tmp);
      if (rc == OXR_OK) {
        e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, s_in, s_out, (int)g, 0,
                                               64, s->s);
      }
    }
    if (rc == OXR_OK && e != hipSuccess) rc = fail_hip(e, "radix sort of group keys");
  } else if (rc == OXR_OK && g == 1) {
    hipError_t e = hipMemcpyAsync(k_out, k_in, 8, hipMemcpyDeviceToDevice, s->s);
    if (e == hipSuccess) e = hipMemcpyAsync(s_out, s_in, 8, hipMemcpyDeviceToDevice, s->s);
    if (e != hipSuccess) rc = fail_hip(e, "hipMemcpyAsync");
  }
  if (rc == OXR_OK) {
    hipError_t e = hipStreamSynchronize(s->s);
    if (e != hipSuccess) rc = fail_hip(e, "aggregate finish");
    else std::memcpy(seen, pin_seen, 8);
  }
  const bool global_only = st->key_col == -1;
  const int has_null = seen[0] ? 1 : 0;
  const int has_min = seen[1] ? 1 : 0;
  uint64_t total = g + has_null + has_min;
  if (global_only) total = 1;  // one row even over no input (SQL global aggregate)
  std::unique_ptr<oxr_batch> ob(new_batch(s, total));
  auto cleanup = [&]() {
    for (auto &c : ob->cols) {
      stream_release(s, c.values);
      stream_release(s, c.validity);
    }
    stream_release(s, tmp);
    stream_release(s, buf);
  };
  if (rc != OXR_OK) {
    cleanup();
    return rc;
  }
  void *key_vals = nullptr;
  uint32_t *key_valid = nullptr;
  if (!global_only) {
    DevCol kc;
    rc = alloc_column(s, st->key_type, 0, total, true, &kc);
    if (rc != OXR_OK) {
      cleanup();
      return rc;
    }
    ob->cols.push_back(kc);
    ob->names.push_back(st->key_name);
    key_vals = kc.values;
    key_valid = (uint32_t *)kc.validity;
  }
  std::vector<OutCol> outs(st->aggs.size());
  for (size_t a = 0; a < st->aggs.size(); a++) {
    DevCol c;
    const bool nullable = !(st->aggs[a].kind == A_COUNT || st->aggs[a].kind == A_COUNT_STAR);
    rc = alloc_column(s, st->out_types[a], 0, total, nullable, &c);
    if (rc != OXR_OK) {
      cleanup();
      return rc;
    }
    ob->cols.push_back(c);
    ob->names.push_back(st->names[a]);
    outs[a].values = c.values;
    outs[a].valid = (uint32_t *)c.validity;
    outs[a].kind = st->aggs[a].kind;
    outs[a].in_type = st->aggs[a].in_type;
    outs[a].out_type = (uint8_t)st->out_types[a];
    outs[a].nullable = nullable;
  }
  void *douts = nullptr;
  rc = stream_alloc(s, sizeof(OutCol) * std::max<size_t>(1, outs.size()), &douts);
  if (rc == OXR_OK && !outs.empty()) {
    hipError_t e = hipMemcpyAsync(douts, outs.data(), sizeof(OutCol) * outs.size(),
                                  hipMemcpyHostToDevice, s->s);
    if (e != hipSuccess) rc = fail_hip(e, "hipMemcpyAsync(out columns)");
  }
  if (rc == OXR_OK && total) {
    Table tt = t;
    int hm = has_min, hn = has_null;
    uint64_t go = g;
    if (global_only) {
      // the one group sits in the NULL-key slot (key_valid false routes there)
      hm = 0;
      hn = 1;
      go = 0;
      // a global aggregate over no rows: COUNT 0, the rest NULL (slot untouched)
    }
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s->s, tt,
                       k_out, s_out, go, hm, hn, st->key_type, key_vals, key_valid,
                       (OutCol *)douts, (int)outs.size(), total, global_only && !seen[0] ? 1 : 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "gather_kernel launch");
  }
  if (rc == OXR_OK) {
    hipError_t e = hipStreamSynchronize(s->s);
    if (e != hipSuccess) rc = fail_hip(e, "aggregate gather");
  }
  stream_release(s, douts);
  if (rc != OXR_OK) {
    cleanup();
    return rc;
  }
  stream_release(s, tmp);
  stream_release(s, buf);
  *out = ob.release();
  return OXR_OK;
}

}  // namespace
}  // namespace oxr

extern "C" {

int oxr_agg_create(oxr_stream *s, const oxr_batch *like, const oxr_agg_spec *spec,
                   oxr_agg_state **out) {
  clear_error();
  if (!s || !like || !spec || !out || (spec->naggs > 0 && !spec->aggs))
    return fail(OXR_EPLAN, "rocm.aggregate", "null argument");
  std::unique_ptr<oxr_agg_state> st(new oxr_agg_state());
  st->dev = s->dev;
  st->owner = retain(s->core);
  OXR_TRY(compile_spec(like, spec, st.get()));
  if (st->hint) OXR_TRY(ensure_capacity(s, st.get(), st->hint));
  *out = st.release();
  return OXR_OK;
}

int oxr_agg_update(oxr_stream *s, oxr_agg_state *st, const oxr_batch *input) {
  clear_error();
  if (!s || !st || !input) return fail(OXR_EPLAN, "rocm.aggregate", "null argument");
  if (st->key_col == -2)
    return fail(OXR_EPLAN, "rocm.aggregate", "a state takes either plain or joined updates");
  return run_update(s, st, input);
}

int oxr_agg_update_join(oxr_stream *s, oxr_agg_state *st, const oxr_join_table *table,
                        const oxr_batch *build, int32_t build_group, const oxr_batch *probe,
                        int32_t probe_key) {
  clear_error();
  if (!s || !st || !table || !build || !probe) return fail(OXR_EPLAN, "rocm.aggregate", "null argument");
  if (table->dev != s->dev) return fail(OXR_EDEVICE, "rocm.aggregate", "join table of another device");
  if (table->build_rows != build->rows)
    return fail(OXR_EPLAN, "rocm.aggregate", "the build batch is not the one the table was built over");
  if (build_group < 0 || build_group >= (int)build->cols.size())
    return fail(OXR_EPLAN, "rocm.aggregate", "build group column out of range");
  const int32_t gt = build->cols[build_group].type;
  if (gt != OXR_INT64 && gt != OXR_INT32)
    return fail(OXR_EUNSUPPORTED, "cpu.int64_column", "group key column has a type other than Int64/Int32");
  if (probe_key < 0 || probe_key >= (int)probe->cols.size())
    return fail(OXR_EPLAN, "rocm.aggregate", "probe key column out of range");
  const int32_t pt = probe->cols[probe_key].type;
  if (pt != OXR_INT64 && pt != OXR_INT32)
    return fail(OXR_EUNSUPPORTED, "cpu.int64_column", "join key column has a type other than Int64/Int32");
  // the groups are the build side's: the state's own key (if any) is unused
  if (st->groups == 0 && st->key_col != -2) {
    st->key_type = gt;
    st->key_name = build->names[build_group];
  } else if (st->key_col != -2 || st->key_type != gt) {
    return fail(OXR_EPLAN, "rocm.aggregate", "a state takes either plain or joined updates, of one key type");
  }
  st->key_col = -2;  // joined: the key column comes from the build side
  JoinSide js;
  js.table = table;
  js.build = build;
  js.build_group = build_group;
  js.probe_slot = st->prog.slot_of(probe_key);
  js.probe_key = probe_key;
  if (js.probe_slot < 0)
    return fail(OXR_EUNSUPPORTED, "rocm.aggregate",
                "an operator reads at most " + std::to_string(kMaxSlots) + " distinct columns");
  return run_update(s, st, probe, &js);
}

int oxr_agg_finish(oxr_stream *s, oxr_agg_state *st, oxr_batch **out) {
  clear_error();
  if (!s || !st || !out) return fail(OXR_EPLAN, "rocm.aggregate", "null argument");
  return run_finish(s, st, out);
}

void oxr_agg_free(oxr_agg_state *st) {
  if (!st) return;
  if (st->mem) st->dev->arena->release(st->mem, st->owner->s);
  release(st->owner);
  delete st;
}

int oxr_aggregate(oxr_stream *s, const oxr_batch *input, const oxr_agg_spec *spec,
                  oxr_batch **out) {
  clear_error();
  if (!s || !input || !spec || !out) return fail(OXR_EPLAN, "rocm.aggregate", "null argument");
  oxr_agg_state *st = nullptr;
  OXR_TRY(oxr_agg_create(s, input, spec, &st));
  std::unique_ptr<oxr_agg_state, void (*)(oxr_agg_state *)> guard(st, oxr_agg_free);
  OXR_TRY(run_update(s, st, input));
  return run_finish(s, st, out);
}

}  // extern "C"

namespace {
int kernel_source(const oxr_column *cols, int32_t ncols, const oxr_agg_spec *spec, int32_t probe_key, bool batched,
                  int32_t compile, char *buf, size_t cap, size_t *len) {
  clear_error();
  if (!spec || ncols < 0 || (ncols > 0 && !cols) || (spec->naggs > 0 && !spec->aggs))
    return fail(OXR_EPLAN, "rocm.jit", "null argument");
  oxr_batch like;
  for (int i = 0; i < ncols; i++) {
    DevCol c;
    c.type = cols[i].type;
    c.dim = cols[i].dim;
    c.validity = cols[i].validity;  // only NULL or not matters
    like.cols.push_back(c);
    like.names.push_back("c" + std::to_string(i));
  }
  oxr_agg_state st;
  OXR_TRY(compile_spec(&like, spec, &st));
  if (probe_key >= ncols) return fail(OXR_EPLAN, "rocm.jit", "probe key column out of range");
  if (batched && (st.key_col != -1 || spec->naggs < 1))
    return fail(OXR_EUNSUPPORTED, "rocm.aggregate_batch", "batched requests have no group key and some aggregate");
  ShapeOf so(&st, probe_key >= 0 && !batched ? probe_key : -2);
  if (batched) so.shape.multi = multi_max_requests((int)st.aggs.size());
  std::string src;
  const int rc = jit_aggregate_source(like.cols, so.shape, compile != 0, &src);
  if (rc == OXR_EUNSUPPORTED)
    return fail(OXR_EUNSUPPORTED, "rocm.jit", "outside what generated kernels cover (the interpreter runs it)");
  OXR_TRY(rc);
  if (len) *len = src.size();
  if (buf && cap) {
    const size_t n = std::min(cap - 1, src.size());
    std::memcpy(buf, src.data(), n);
    buf[n] = 0;
  }
  return OXR_OK;
}
}  // namespace

extern "C" int oxr_aggregate_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_agg_spec *spec,
                                           int32_t probe_key, int32_t compile, char *buf, size_t cap,
                                           size_t *len) {
  return kernel_source(cols, ncols, spec, probe_key, false, compile, buf, cap, len);
}

extern "C" int oxr_agg_batch_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_agg_spec *spec,
                                           int32_t compile, char *buf, size_t cap, size_t *len) {
  return kernel_source(cols, ncols, spec, -1, true, compile, buf, cap, len);
}

// ---------------------------------------------------------- batched requests
// Many scalar aggregations of one prepared shape over one resident table,
// answered by one pass over the table (aggregate_multi, aggregate_jit.h).
// The chain of one run -- the requests' literals copied from pinned host
// memory, the shared scan, the final fold, the result records copied back
// to pinned host memory -- is captured once per batch size into a hipGraph
// and replayed for every run; nothing in it is a memset, and no device
// memory is allocated per run.
struct oxr_agg_batch {
  oxr_device *dev = nullptr;
  StreamCore *owner = nullptr;
  std::mutex mu;  // runs take turns on the buffers below
  std::unique_ptr<oxr_agg_state> proto;  // the prepared spec
  std::vector<DevCol> cols;              // the table's columns (views)
  std::shared_ptr<void> hold;            // keeps the table's buffers alive
  uint64_t rows = 0;
  std::string source;                    // the shape's generated source
  JitAgg ja;
  int kmax = 1;                          // requests per run
  uint32_t nlit = 0;                     // literals per request
  unsigned grid = 1;
  AggJitArgs args{};
  void *dmem = nullptr;                  // lits | partials | results (arena)
  uint64_t *dlits = nullptr;
  int64_t *dpart = nullptr, *dres = nullptr;
  uint64_t *hlits = nullptr;             // pinned
  int64_t *hres = nullptr;               // pinned
  hipStream_t cap = nullptr;             // capture-only stream
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace oxr {
namespace {

// The generated shape of `spec` over the batch's table and its literals.
int batch_shape(const oxr_agg_batch *b, const oxr_agg_spec *spec, std::string *src, std::vector<uint64_t> *lits) {
  oxr_batch like;
  like.rows = 0;
  like.cols = b->cols;
  like.names.assign(b->cols.size(), std::string());
  oxr_agg_state st;
  OXR_TRY(compile_spec(&like, spec, &st));
  if (st.key_col != -1) return fail(OXR_EPLAN, "rocm.aggregate_batch", "a request with a group key");
  ShapeOf so(&st, -2);
  so.shape.multi = b->kmax;
  JitAgg ja;
  const int rc = jit_aggregate_generate(b->cols, so.shape, &ja);
  if (rc == OXR_EUNSUPPORTED)
    return fail(OXR_EUNSUPPORTED, "rocm.aggregate_batch", "the shape is outside the generated kernels");
  OXR_TRY(rc);
  *src = std::move(ja.source);
  *lits = std::move(ja.lits);
  return OXR_OK;
}

int batch_bucket(int n) {
  for (int i = 0; i < 4; i++)
    if (kMultiK[i] >= n) return i;
  return 3;
}

// The chain of one run of bucket `ki` on `st`.
int batch_enqueue(oxr_agg_batch *b, hipStream_t st, int ki) {
  const int K = kMultiK[ki];
  const int na = (int)b->proto->aggs.size();
  OXR_HIP(hipMemcpyAsync(b->dlits, b->hlits, (1 + (size_t)K * b->nlit) * 8, hipMemcpyHostToDevice, st));
  void *params[] = {&b->args};
  OXR_HIP(hipModuleLaunchKernel(b->ja.multi_fn[ki], b->grid, 1, 1, 256, 1, 1, 0, st, params, nullptr));
  OXR_HIP(hipModuleLaunchKernel(b->ja.multi_final[ki], (unsigned)K, 1, 1, 256, 1, 1, 0, st, params, nullptr));
  OXR_HIP(hipMemcpyAsync(b->hres, b->dres, (size_t)K * (na + 1) * 8, hipMemcpyDeviceToHost, st));
  return OXR_OK;
}

int batch_graph(oxr_agg_batch *b, int ki) {
  if (b->exec[ki]) return OXR_OK;
  if (!b->cap) OXR_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  OXR_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  const int rc = batch_enqueue(b, b->cap, ki);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (rc != OXR_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail_hip(e, "capturing the batched aggregation");
  const hipError_t ei = hipGraphInstantiate(&b->exec[ki], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) {
    b->exec[ki] = nullptr;
    return fail_hip(ei, "instantiating the batched aggregation graph");
  }
  return OXR_OK;
}

void batch_destroy(oxr_agg_batch *b) {
  for (auto &x : b->exec)
    if (x) hipGraphExecDestroy(x);
  if (b->cap) hipStreamDestroy(b->cap);
  if (b->hlits) hipHostFree(b->hlits);
  if (b->hres) hipHostFree(b->hres);
  if (b->dmem) b->dev->arena->release(b->dmem, b->owner->s);
  if (b->owner) release(b->owner);
  delete b;
}

}  // namespace
}  // namespace oxr

extern "C" {

int oxr_agg_batch_prepare(oxr_stream *s, const oxr_batch *table, const oxr_agg_spec *spec,
                          int32_t *max_requests, oxr_agg_batch **out) {
  clear_error();
  if (!s || !table || !spec || !out || (spec->naggs > 0 && !spec->aggs))
    return fail(OXR_EPLAN, "rocm.aggregate_batch", "null argument");
  if (spec->group_by != -1) return fail(OXR_EUNSUPPORTED, "rocm.aggregate_batch", "batched requests have no group key");
  if (spec->naggs < 1) return fail(OXR_EPLAN, "rocm.aggregate_batch", "no aggregates");
  std::unique_ptr<oxr_agg_batch, void (*)(oxr_agg_batch *)> b(new oxr_agg_batch(), batch_destroy);
  b->dev = s->dev;
  b->owner = retain(s->core);
  b->proto.reset(new oxr_agg_state());
  OXR_TRY(compile_spec(table, spec, b->proto.get()));
  b->cols = table->cols;
  for (auto &c : b->cols) c.owned = false;
  b->hold = share_columns(table);
  b->rows = table->rows;
  const int na = (int)b->proto->aggs.size();
  b->kmax = multi_max_requests(na);
  {
    ShapeOf so(b->proto.get(), -2);
    so.shape.multi = b->kmax;
    const int rc = jit_aggregate(s->dev, b->cols, so.shape, &b->ja);
    if (rc == OXR_EUNSUPPORTED)
      return fail(OXR_EUNSUPPORTED, "rocm.aggregate_batch", "the shape is outside the generated kernels");
    OXR_TRY(rc);
  }
  b->source = b->ja.source;
  b->nlit = (uint32_t)b->ja.lits.size();
  // the grid: a fixed function of the table and the device, so a request's
  // order of additions never changes
  const uint64_t tile = 256ull * (uint64_t)b->ja.multi_rows;
  const uint64_t ntiles = std::max<uint64_t>(1, (b->rows + tile - 1) / tile);
  const uint64_t per_cu = (uint64_t)std::max<int64_t>(1, tune_get("agg.multi_grid", 2));
  b->grid = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)std::max(1, s->dev->cu_count) * per_cu);
  const int K = b->kmax;
  const size_t lit_b = ((1 + (size_t)K * b->nlit) * 8 + 255) / 256 * 256;
  const size_t part_b = ((size_t)b->grid * K * (2 * na + 1) * 8 + 255) / 256 * 256;
  const size_t res_b = (size_t)K * (na + 1) * 8;
  OXR_TRY(stream_alloc(s, lit_b + part_b + res_b, &b->dmem));
  b->dlits = (uint64_t *)b->dmem;
  b->dpart = (int64_t *)((char *)b->dmem + lit_b);
  b->dres = (int64_t *)((char *)b->dmem + lit_b + part_b);
  OXR_HIP(hipHostMalloc((void **)&b->hlits, lit_b, hipHostMallocDefault));
  OXR_HIP(hipHostMalloc((void **)&b->hres, res_b, hipHostMallocDefault));
  std::memset(&b->args, 0, sizeof b->args);
  b->args.n = b->rows;
  for (size_t i = 0; i < b->cols.size() && i < (size_t)kAggCols; i++) {
    b->args.col[i] = b->cols[i].values;
    b->args.valid[i] = (const uint32_t *)b->cols[i].validity;
  }
  b->args.mlits = b->dlits;
  b->args.mnlit = b->nlit;
  b->args.mblocks = b->grid;
  b->args.mpart = b->dpart;
  b->args.mres = b->dres;
  // every run waits for the table's producer: the table was made on `s`
  OXR_HIP(hipStreamSynchronize(s->s));
  if (max_requests) *max_requests = K;
  *out = b.release();
  return OXR_OK;
}

int oxr_agg_batch_columns(const oxr_agg_batch *b, int32_t *naggs, int32_t *types, const char **names) {
  clear_error();
  if (!b || !naggs) return fail(OXR_EPLAN, "rocm.aggregate_batch", "null argument");
  const int na = (int)b->proto->aggs.size();
  *naggs = na;
  for (int a = 0; a < na; a++) {
    if (types) types[a] = b->proto->out_types[a];
    if (names) names[a] = b->proto->names[a].c_str();
  }
  return OXR_OK;
}

int oxr_agg_batch_run(oxr_stream *s, oxr_agg_batch *b, const oxr_agg_spec *requests, int32_t n,
                      int64_t *values, uint8_t *valid, int32_t *status) {
  clear_error();
  if (!s || !b || (n > 0 && (!requests || !values || !valid || !status)))
    return fail(OXR_EPLAN, "rocm.aggregate_batch", "null argument");
  if (s->dev != b->dev) return fail(OXR_EPLAN, "rocm.aggregate_batch", "the stream is on another device");
  if (n < 0 || n > b->kmax)
    return fail(OXR_EPLAN, "rocm.aggregate_batch",
                "a run takes 0 to " + std::to_string(b->kmax) + " requests, not " + std::to_string(n));
  if (n == 0) return OXR_OK;
  const int na = (int)b->proto->aggs.size();
  // every request must be the prepared shape; its literals fill its row
  std::vector<uint64_t> lits((size_t)n * b->nlit);
  for (int r = 0; r < n; r++) {
    std::string src;
    std::vector<uint64_t> l;
    OXR_TRY(batch_shape(b, &requests[r], &src, &l));
    if (src != b->source || l.size() != b->nlit)
      return fail(OXR_EPLAN, "rocm.aggregate_batch",
                  "request " + std::to_string(r) + " is not the prepared shape (only literal values may differ)");
    std::copy(l.begin(), l.end(), lits.begin() + (size_t)r * b->nlit);
  }
  const int ki = batch_bucket(n);
  std::lock_guard<std::mutex> lock(b->mu);
  b->hlits[0] = (uint64_t)n;
  std::copy(lits.begin(), lits.end(), b->hlits + 1);
  // the kernel runs all of its K request slots: the unused ones repeat the
  // last request's literals (their results are never read)
  for (int r = n; r < kMultiK[ki]; r++)
    std::copy(lits.end() - b->nlit, lits.end(), b->hlits + 1 + (size_t)r * b->nlit);
  if (tune_get("agg.batch_graph", 1) > 0) {
    OXR_TRY(batch_graph(b, ki));
    ProfScope prof(s->s, "aggregate_batch");
    OXR_HIP(hipGraphLaunch(b->exec[ki], s->s));
  } else {
    ProfScope prof(s->s, "aggregate_batch");
    OXR_TRY(batch_enqueue(b, s->s, ki));
  }
  OXR_HIP(hipStreamSynchronize(s->s));
  for (int r = 0; r < n; r++) {
    const int64_t *rec = b->hres + (size_t)r * (na + 1);
    const uint32_t fl = (uint32_t)rec[na];
    status[r] = (fl & kMultiErr) ? OXR_EEXECUTION : OXR_OK;
    for (int a = 0; a < na; a++) {
      const uint8_t k = b->proto->aggs[a].kind;
      const bool cnt = k == A_COUNT || k == A_COUNT_STAR;
      values[(size_t)r * na + a] = rec[a];
      valid[(size_t)r * na + a] = (uint8_t)(cnt || ((fl >> a) & 1u));
    }
  }
  return OXR_OK;
}

void oxr_agg_batch_free(oxr_agg_batch *b) {
  if (b) batch_destroy(b);
}

}  // extern "C"

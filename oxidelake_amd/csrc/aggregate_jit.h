// Grouped aggregation on the device: the HBM group table, the accumulators,
// and the skeleton of the aggregation kernels that jit.hip generates per
// aggregation shape and compiles at run time with hipRTC. Shared with the
// interpreter kernel in aggregate.hip; compiles under hipcc and hipRTC.
//
// Semantics (crates/oxidelake-device/src/cpu/kernels.rs:478-818): a NULL key
// is a group of its own; SUM/MIN/MAX skip NULL inputs; COUNT counts non-NULL
// inputs; SUM over integers wraps in Int64; SUM over floats is an FP64 sum
// from +0.0; MIN/MAX of floats compare totalOrder keys (kernels.rs:820-838).
//
// FP64 sums are plain atomic sums by default. In exact mode (knob
// agg.exact_f64) they are compensated: every accumulator (lane pair, LDS
// slot, HBM slot) is a pair (hi, lo) whose lo gathers the exact rounding
// error of each addition into hi (TwoSum, Knuth TAOCP vol. 2, 4.2.2), so the
// result hi + lo is the exact sum up to second-order terms, in whatever order
// the atomics land: the same from run to run even where the sum cancels, at
// the price of a second atomic per update.
#pragma once

#include "jtable.h"

namespace oxr {

constexpr int kMaxAggs = 8;
constexpr int kMaxSpillParts = 1024;  // partitions of the spill range
constexpr int kSpillBuckets = 32;     // coarse partitions a block appends to
constexpr int kLdsLog2 = 10;
constexpr int kLdsSlots = 1 << kLdsLog2;  // block table slots
constexpr int kLdsAggs = 4;           // aggregates pre-aggregated in LDS
constexpr int kLdsProbe = 8;          // LDS probe limit before going global
constexpr int64_t kEmpty = kI64Min;   // empty-slot marker
constexpr int64_t kI64Max = (int64_t)(~(uint64_t)0 >> 1);

enum : uint8_t {
  A_COUNT = 1,
  A_SUM_I64,
  A_SUM_F64,
  A_MIN_I64,
  A_MAX_I64,
  A_MIN_F64,  // over totalOrder keys of f64
  A_MAX_F64,
  A_MIN_F32,  // over totalOrder keys of f32 (sign-extended)
  A_MAX_F32,
  A_COUNT_STAR
};

__host__ __device__ inline int64_t acc_init(uint8_t kind) {
  switch (kind) {
    case A_MIN_I64:
    case A_MIN_F64:
    case A_MIN_F32: return kI64Max;
    case A_MAX_I64:
    case A_MAX_F64:
    case A_MAX_F32: return kI64Min;
  }
  return 0;
}

__host__ __device__ inline uint64_t fmix64(uint64_t h) { return join_hash(h); }

// Block-table slot of a key: Fibonacci hashing of the key's two halves
// folded together (3 instructions against fmix64's ~16). The block table
// is a cache in front of the HBM table, so a weaker hash only costs probes
// (a key it cannot place goes to HBM), never results.
__device__ inline uint32_t lds_hash(int64_t key) {
  const uint32_t f = (uint32_t)key ^ (uint32_t)((uint64_t)key >> 32);
  return (f * 0x9E3779B1u) >> (32 - kLdsLog2);
}

// The HBM group table: open addressing, one 64-bit CAS claims a key.
struct Table {
  uint64_t cap;       // power of two; slots cap (NULL key) and cap+1 (INT64_MIN key) follow
  int64_t *keys;      // cap+2
  int64_t *acc[kMaxAggs];
  int64_t *lo[kMaxAggs];    // FP64 SUM: the rounding errors of acc (doubles), else NULL
  uint32_t *has[kMaxAggs];  // cap+2 each, or NULL
  uint32_t *seen;     // [0] NULL-key group, [1] INT64_MIN group
  uint64_t *distinct; // new keys claimed
  uint32_t *overflow; // a key found no slot
};

__device__ inline void acc_apply_global(uint8_t kind, int64_t *addr, int64_t v) {
  switch (kind) {
    case A_COUNT:
    case A_COUNT_STAR:
    case A_SUM_I64:
      atomicAdd((unsigned long long *)addr, (unsigned long long)v);
      break;
    case A_SUM_F64:
      unsafeAtomicAdd((double *)addr, as_f64((uint64_t)v));
      break;
    case A_MIN_I64:
    case A_MIN_F64:
    case A_MIN_F32:
      atomicMin((long long *)addr, (long long)v);
      break;
    default:
      atomicMax((long long *)addr, (long long)v);
      break;
  }
}

__device__ inline void acc_apply_lds(uint8_t kind, int64_t *addr, int64_t v) {
  switch (kind) {
    case A_COUNT:
    case A_COUNT_STAR:
    case A_SUM_I64:
      atomicAdd((unsigned long long *)addr, (unsigned long long)v);
      break;
    case A_SUM_F64:
      unsafeAtomicAdd((double *)addr, as_f64((uint64_t)v));
      break;
    case A_MIN_I64:
    case A_MIN_F64:
    case A_MIN_F32:
      atomicMin((long long *)addr, (long long)v);
      break;
    default:
      atomicMax((long long *)addr, (long long)v);
      break;
  }
}

// The exact error of s = fl(a + b) for a finite s (TwoSum; needs the adds
// compiled without contraction).
__device__ inline double two_sum_err(double a, double b, double s) {
  const double bb = s - a;
  return (a - (s - bb)) + (b - bb);
}

// Adds the compensated value (v, vlo) to the accumulator (hi, lo). The
// atomic returns the value it added v to, so the rounding error of that
// addition is recovered exactly and goes to lo with vlo. Works on LDS and
// HBM addresses alike.
__device__ inline void sum_f64(int64_t *hi, int64_t *lo, double v, double vlo) {
  const double old = unsafeAtomicAdd((double *)hi, v);
  const double s = old + v;
  double e = vlo;
  if (__builtin_isfinite(s)) e += two_sum_err(old, v, s);
  if (e != 0.0) unsafeAtomicAdd((double *)lo, e);
}

// The compensated sum of (hi, lo) over the wave (every lane gets it).
__device__ inline void wave_sum_f64(double &hi, double &lo) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const double oh = __shfl_xor(hi, d, 64);
    const double ol = __shfl_xor(lo, d, 64);
    const double s = hi + oh;
    const double e = __builtin_isfinite(s) ? two_sum_err(hi, oh, s) : 0.0;
    hi = s;
    lo = (lo + ol) + e;
  }
}

// hi + lo as the accumulator's value (lo == 0 keeps hi's sign of zero).
// A non-finite hi is the sum (inf or NaN absorbs every later addition, and
// lo may then hold the NaN of an unguarded error term).
__host__ __device__ inline double sum_value(double hi, double lo) {
  return lo != 0.0 && __builtin_isfinite(hi) ? hi + lo : hi;
}

// The per-row contribution of a value of `in_type` (widened to 64 bits:
// ints sign-extended, floats as their bits) to an accumulator.
__device__ inline int64_t contribution(uint8_t kind, uint8_t in_type, uint64_t v) {
  switch (kind) {
    case A_COUNT:
    case A_COUNT_STAR: return 1;
    case A_SUM_I64:
    case A_MIN_I64:
    case A_MAX_I64: return (int64_t)v;
    case A_SUM_F64:
      return (int64_t)(in_type == 3 /* OXR_FLOAT32 */ ? f64_bits((double)as_f32(v)) : v);
    case A_MIN_F64:
    case A_MAX_F64: {
      // totalOrder without zero normalisation (kernels.rs:527-538)
      const int64_t b = (int64_t)v;
      return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
    }
    default: {
      const int32_t b = (int32_t)(uint32_t)v;
      return (int64_t)(b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1));
    }
  }
}

// MIN/MAX of floats compare these keys (totalOrder, no zero normalisation).
__device__ inline int64_t total_order64(double x) {
  const int64_t b = (int64_t)f64_bits(x);
  return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}
__device__ inline int64_t total_order32(float x) {
  const int32_t b = (int32_t)(uint32_t)f32_bits(x);
  return (int64_t)(b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1));
}

// Finds (or claims) the HBM slot of `key` (not NULL, not INT64_MIN).
__device__ inline uint64_t global_slot(const Table &t, int64_t key, bool *ok) {
  const uint64_t mask = t.cap - 1;
  uint64_t h = fmix64((uint64_t)key) & mask;
  for (uint64_t probe = 0; probe < t.cap; probe++) {
    const int64_t cur = __hip_atomic_load(&t.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == key) return h;
    if (cur == kEmpty) {
      const unsigned long long old =
          atomicCAS((unsigned long long *)&t.keys[h], (unsigned long long)kEmpty, (unsigned long long)key);
      if ((int64_t)old == kEmpty) {
        atomicAdd((unsigned long long *)t.distinct, 1ull);
        return h;
      }
      if ((int64_t)old == key) return h;
    }
    h = (h + 1) & mask;
  }
  *ok = false;
  return 0;
}

__device__ inline uint64_t slot_of(const Table &t, bool key_valid, int64_t key, bool *ok) {
  if (!key_valid) {
    t.seen[0] = 1;
    return t.cap;
  }
  if (key == kEmpty) {
    t.seen[1] = 1;
    return t.cap + 1;
  }
  return global_slot(t, key, ok);
}

// ------------------------------------------------- generated-kernel skeleton
// The generated part is a struct Q:
//   kAggs, kLdsAggs, kKind[kAggs], kNullable[kAggs], kAnyNullable,
//   kExactF64 (compensated FP64 sums)
//   struct In, ins(p, base); struct Row, load(i, j, in, x)   (as for filters)
//   pred(L, x, j, in, err)              the fused filter (true: keep; L: the
//                                       literals, in prepared mode)
//   key(p, x, j, in, key, kv, hid)      the group key (false: no group, i.e.
//                                       no join match); kv false = NULL key;
//                                       hid: the group's hot id when the join
//                                       table's group entries carry it (-1:
//                                       cold), else -2
//   contrib(L, x, j, sel, c, cm, err)   each aggregate's contribution; bit a
//                                       of cm set where input a is not NULL
constexpr int kAggCols = 16;

struct AggJitArgs {
  uint64_t n;                   // input (probe) rows
  const void *col[kAggCols];
  const uint32_t *valid[kAggCols];
  Table t;
  uint32_t *error;              // [0] division by zero
  // join: the group of each table slot, built once per join (unique keys)
  JTable jt;
  const uint64_t *gos;          // Int32 groups: (state << 32) | group
  const uint64_t *gos64;        // Int64 groups: {group, state} pairs
  uint32_t gos_hot;             // the states carry hot ids (aggregate_hot): state | (hid + 1) << 8
  uint64_t lit[32];             // prepared mode: literal bits
  // dense kernels: keys [dense_min, dense_min + dense_keys), plus the NULL
  // key's slot at index dense_keys when dense_null
  int64_t dense_min;
  uint32_t dense_keys, dense_null;
  // spill: rows the block table does not hold whose key is in [smin, smin +
  // srange) go to `spill` instead of HBM atomics, as records of
  // SpillRec<Q>::kWords 32-bit words (see SpillRec). Block b appends
  // to kSpillBuckets buckets of spill_cap records each (bucket c at record
  // (b * kSpillBuckets + c) * spill_cap; partition p goes to bucket p /
  // sper), so a few tails per block stay in L2 while they fill. Per block:
  // the bucket counts (spill_n[b * kSpillBuckets + c]) and a histogram over
  // partitions of 2^slog2 groups (hist[p * gridDim + b]).
  uint32_t *spill;              // NULL: no spilling
  uint64_t spill_cap;           // records per bucket
  uint32_t *spill_n;
  uint32_t *hist;
  int64_t smin;
  uint64_t srange;
  uint32_t slog2, sparts, sper;
  // hot groups (aggregate_hot): hot id of each group of the spill range
  // (hotmap[key - smin], -1 cold); each block's hot accumulators go to
  // hot_part[(block * kAggs + a) * nhot + h] and its flags to
  // hot_flags[block * nhot + h] (bit 0: a row of the group was selected;
  // bit 1 + a: aggregate a saw a non-NULL input) for the reduce that follows
  const int32_t *hotmap;
  uint32_t nhot;
  int64_t *hot_part;
  uint8_t *hot_flags;
  // batched requests (aggregate_multi): mlits[0] = the requests of this
  // run, request r's literals at mlits[1 + r * mnlit ...]; each of the
  // multi kernel's mblocks blocks writes its partials to mpart, and the
  // final fold writes one result record per request to mres
  const uint64_t *mlits;
  uint32_t mnlit, mblocks;
  int64_t *mpart;
  int64_t *mres;
};

// Hot groups a block accumulates in LDS (aggregate_hot): 40 KB of
// accumulators and flags for NA aggregates, a multiple of 256.
__host__ __device__ constexpr int hot_capacity(int na, bool nullable) {
  return (40960 / (8 * na + 1 + (nullable ? na : 0))) / 256 * 256;
}

// group of a probe key through the join: false when nothing matches. With
// p.gos_hot the group entries also carry the group's hot id (the state
// word's bits 8 and up hold hot id + 1, 0 for a cold group).
__device__ inline bool join_group(const AggJitArgs &p, int64_t pk, int64_t &g, bool &gv, int32_t &hid) {
  uint64_t slot;
  if (p.jt.dense) {
    // every slot of a dense table has an entry (state 0: no build row)
    slot = (uint64_t)pk - (uint64_t)p.jt.min_key;
    if (pk < p.jt.min_key || slot >= p.jt.slots) return false;
  } else {
    slot = join_find(p.jt, pk);
    if (slot == kNoSlot) return false;
  }
  uint32_t state;
  if (p.gos) {
    const uint64_t e = p.gos[slot];
    state = (uint32_t)(e >> 32);
    g = (int64_t)(int32_t)(uint32_t)e;
  } else {
    g = (int64_t)p.gos64[2 * slot];
    state = (uint32_t)p.gos64[2 * slot + 1];
  }
  if (p.gos_hot) hid = (int32_t)(state >> 8) - 1;
  state &= 0xFFu;
  gv = state == 1;  // 2: the build row's group is NULL
  return state != 0;
}

template <uint8_t K>
__device__ inline int64_t combine(int64_t a, int64_t b) {
  if (K == A_SUM_F64) return (int64_t)f64_bits(as_f64((uint64_t)a) + as_f64((uint64_t)b));
  if (K == A_MIN_I64 || K == A_MIN_F64 || K == A_MIN_F32) return a < b ? a : b;
  if (K == A_MAX_I64 || K == A_MAX_F64 || K == A_MAX_F32) return a > b ? a : b;
  return (int64_t)((uint64_t)a + (uint64_t)b);
}

template <uint8_t K>
__device__ inline int64_t wave_combine(int64_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = combine<K>(v, (int64_t)__shfl_xor((long long)v, d, 64));
  return v;
}

// Index of aggregate a among the FP64 sums before it (its LDS lo array).
template <class Q>
__host__ __device__ constexpr int lo_index(int a) {
  int n = 0;
  for (int i = 0; i < a; i++) n += Q::kKind[i] == A_SUM_F64;
  return n;
}

// A spilled row is a record of 32-bit words: the group index (key - smin);
// when any input is nullable, the mask of aggregates whose input was not
// NULL; then aggregate a's contribution in Q::kSpillWords[a] words: 0 for
// COUNT and COUNT(*) (a contribution of 1, under the mask), 1 for a value
// that fits 32 bits (Int32 inputs of SUM / MIN / MAX, Float32 MIN / MAX
// keys; sign-extended when read back), 2 for 64 bits, low word first.
template <class Q>
struct SpillRec {
  static constexpr int kHdr = Q::kAnyNullable ? 2 : 1;
  static constexpr int off(int a) {
    int o = kHdr;
    for (int i = 0; i < a; i++) o += Q::kSpillWords[i];
    return o;
  }
  static constexpr int kWords = off(Q::kAggs);
  __device__ static inline void pack(uint32_t (&w)[kWords], uint32_t g, uint32_t cm, const int64_t *c) {
    w[0] = g;
    if (kHdr == 2) w[1] = cm;
#pragma unroll
    for (int a = 0; a < Q::kAggs; a++) {
      if (Q::kSpillWords[a] >= 1) w[off(a)] = (uint32_t)(uint64_t)c[a];
      if (Q::kSpillWords[a] == 2) w[off(a) + 1] = (uint32_t)((uint64_t)c[a] >> 32);
    }
  }
  __device__ static inline int64_t value(const uint32_t (&w)[kWords], int a) {
    if (Q::kSpillWords[a] == 0) return 1;
    if (Q::kSpillWords[a] == 1) return (int64_t)(int32_t)w[off(a)];
    return (int64_t)((uint64_t)w[off(a)] | ((uint64_t)w[off(a) + 1] << 32));
  }
  __device__ static inline uint32_t mask(const uint32_t (&w)[kWords]) {
    return kHdr == 2 ? w[1] : (1u << Q::kAggs) - 1u;
  }
};

// R consecutive 32-bit words at dst / src (aligned to 4 * R bytes when R is
// even), in 16- or 8-byte accesses where R allows.
template <int R>
__device__ __forceinline__ void store_words(uint32_t *dst, const uint32_t (&w)[R]) {
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int i = 0; i < R / 4; i++) ((uint4 *)dst)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
  } else if constexpr (R % 2 == 0) {
#pragma unroll
    for (int i = 0; i < R / 2; i++) ((uint2 *)dst)[i] = make_uint2(w[2 * i], w[2 * i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < R; i++) dst[i] = w[i];
  }
}
typedef uint32_t u32x4_nt __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_nt __attribute__((ext_vector_type(2)));
// A record's words read once: nontemporal loads.
template <int R>
__device__ __forceinline__ void load_words_nt(const uint32_t *src, uint32_t (&w)[R]) {
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int i = 0; i < R / 4; i++) {
      const u32x4_nt v = __builtin_nontemporal_load((const u32x4_nt *)src + i);
      w[4 * i] = v.x;
      w[4 * i + 1] = v.y;
      w[4 * i + 2] = v.z;
      w[4 * i + 3] = v.w;
    }
  } else if constexpr (R % 2 == 0) {
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      const u32x2_nt v = __builtin_nontemporal_load((const u32x2_nt *)src + i);
      w[2 * i] = v.x;
      w[2 * i + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; i++) w[i] = __builtin_nontemporal_load(src + i);
  }
}
template <int R>
__device__ __forceinline__ void load_words(const uint32_t *src, uint32_t (&w)[R]) {
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int i = 0; i < R / 4; i++) {
      const uint4 v = ((const uint4 *)src)[i];
      w[4 * i] = v.x;
      w[4 * i + 1] = v.y;
      w[4 * i + 2] = v.z;
      w[4 * i + 3] = v.w;
    }
  } else if constexpr (R % 2 == 0) {
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      const uint2 v = ((const uint2 *)src)[i];
      w[2 * i] = v.x;
      w[2 * i + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; i++) w[i] = src[i];
  }
}

// A selected row (key, kv, contributions c under mask cm) that the block's
// LDS does not hold: its spill bucket when its key is in the spill range and
// the bucket has room (true), else the caller applies it to the HBM table.
template <class Q>
__device__ __forceinline__ bool spill_row(const AggJitArgs &p, uint32_t *spill_blk, uint32_t *s_bucket,
                                          uint32_t *s_hist, int64_t key, bool kv, const int64_t *c,
                                          uint32_t cm) {
  using SR = SpillRec<Q>;
  if (!(p.spill && kv && key >= p.smin && (uint64_t)key - (uint64_t)p.smin < p.srange)) return false;
  const uint32_t g = (uint32_t)((uint64_t)key - (uint64_t)p.smin);
  const uint32_t part = g >> p.slog2, bucket = part / p.sper;
  const uint32_t pos = atomicAdd(&s_bucket[bucket], 1u);
  if (pos >= p.spill_cap) return false;
  // 32-bit offset inside the block's buckets (the host keeps
  // kSpillBuckets * spill_cap * kWords below 2^31)
  uint32_t w[SR::kWords];
  SR::pack(w, g, cm, c);
  store_words<SR::kWords>(spill_blk + (bucket * (uint32_t)p.spill_cap + pos) * (uint32_t)SR::kWords, w);
  atomicAdd(&s_hist[part], 1u);
  return true;
}

// The spill's block region: block b's kSpillBuckets buckets of spill_cap
// records.
template <class Q>
__device__ __forceinline__ uint32_t *spill_block(const AggJitArgs &p) {
  return p.spill ? p.spill + (uint64_t)blockIdx.x * kSpillBuckets * p.spill_cap * SpillRec<Q>::kWords : nullptr;
}

// log2 of the groups one fold workgroup holds in LDS (8-byte accumulators,
// a used byte, and a non-NULL byte per nullable aggregate), within 64 KB.
__host__ __device__ constexpr uint32_t fold_log2(int na, int nnull) {
  uint32_t l = 14;
  while (l > 5 && ((uint64_t)1 << l) * (uint64_t)(8 * na + 1 + nnull) > 65536) l--;
  return l;
}

// Arguments of the spill's scatter and fold kernels (see aggregate.hip's
// spill passes): the block buckets (spill, spill_n, spill_cap), the
// exclusive scan of the histogram (offs[part * B + block] = where block's
// records of partition part start in recs), and the partitioned records.
struct SpillArgs {
  const uint32_t *spill;
  const uint32_t *spill_n;
  const uint32_t *offs;
  uint32_t *recs;
  uint64_t cap;
  uint32_t B, nparts, sper, slog2;
  int64_t smin;
  Table t;
  uint32_t *error;  // [1]: the fold's LDS does not hold 2^slog2 groups
};

// Workgroup w = b * kSpillBuckets + c moves block b's bucket c to its fine
// partitions' ranges (bucket c holds partitions [c * sper, (c + 1) * sper));
// each thread moves 4 records per step, loads first.
template <class Q>
__device__ __forceinline__ void spill_scatter(const SpillArgs &a) {
  constexpr int R = SpillRec<Q>::kWords;
  __shared__ uint32_t s_cur[kMaxSpillParts];
  const uint32_t w = blockIdx.x, b = w / kSpillBuckets, c = w % kSpillBuckets;
  const uint32_t p0 = c * a.sper;
  for (uint32_t f = threadIdx.x; f < a.sper && p0 + f < a.nparts; f += 256) s_cur[f] = a.offs[(uint64_t)(p0 + f) * a.B + b];
  __syncthreads();
  const uint64_t n = a.spill_n[w] < a.cap ? a.spill_n[w] : a.cap;
  const uint32_t *src = a.spill + (uint64_t)w * a.cap * R;
  for (uint64_t r0 = threadIdx.x; r0 < n; r0 += 4 * 256) {
    uint32_t rec[4][R];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint64_t r = r0 + k * 256;
      if (r < n) {
        if constexpr (Q::kSpillNT) load_words_nt<R>(src + r * R, rec[k]);
        else load_words<R>(src + r * R, rec[k]);
      }
      else
#pragma unroll
        for (int x = 0; x < R; x++) rec[k][x] = 0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (r0 + k * 256 >= n) continue;
      const uint32_t pos = atomicAdd(&s_cur[(rec[k][0] >> a.slog2) - p0], 1u);
      store_words<R>(a.recs + (uint64_t)pos * R, rec[k]);
    }
  }
}

// Index of aggregate a among the nullable ones before it.
template <class Q>
__host__ __device__ constexpr int null_index(int a) {
  int n = 0;
  for (int i = 0; i < a; i++) n += Q::kNullable[i] != 0;
  return n;
}

// Workgroup p folds partition p's records into an LDS table of 2^slog2
// groups (accumulators, used and non-NULL flags: plain byte stores), then
// adds each used group's accumulators to the HBM table. kFoldThreads
// threads each keep kFoldRecs records' loads in flight (a partition is a
// few MB of records and there are about two partitions per CU: 256 threads
// x 4 records left the fold waiting on memory).
constexpr int kFoldThreads = 512, kFoldRecs = 8;
template <class Q>
__device__ __forceinline__ void spill_fold(const SpillArgs &sa) {
  using SR = SpillRec<Q>;
  constexpr int NA = Q::kAggs, R = SR::kWords, NN = null_index<Q>(NA);
  constexpr uint32_t kG = 1u << fold_log2(NA, NN);
  __shared__ int64_t s_acc[NA][kG];
  __shared__ uint8_t s_used[kG];
  __shared__ uint8_t s_has[NN > 0 ? NN : 1][NN > 0 ? kG : 1];
  const uint32_t G = 1u << sa.slog2;
  if (G > kG) {  // the host sizes partitions with fold_log2 too
    if (threadIdx.x == 0) atomicOr(&sa.error[1], 1u);
    return;
  }
  for (uint32_t i = threadIdx.x; i < G; i += kFoldThreads) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      s_acc[a][i] = acc_init(Q::kKind[a]);
      if (Q::kNullable[a]) s_has[null_index<Q>(a)][i] = 0;
    }
    s_used[i] = 0;
  }
  __syncthreads();
  const uint32_t part = blockIdx.x;
  const uint64_t lo = sa.offs[(uint64_t)part * sa.B], hi = sa.offs[(uint64_t)(part + 1) * sa.B];
  for (uint64_t r0 = lo + threadIdx.x; r0 < hi; r0 += (uint64_t)kFoldRecs * kFoldThreads) {
    uint32_t rec[kFoldRecs][R];
#pragma unroll
    for (int k = 0; k < kFoldRecs; k++) {
      const uint64_t r = r0 + (uint64_t)k * kFoldThreads;
      if (r < hi) {
        if constexpr (Q::kSpillNT) load_words_nt<R>(sa.recs + r * R, rec[k]);
        else load_words<R>(sa.recs + r * R, rec[k]);
      }
      else
#pragma unroll
        for (int x = 0; x < R; x++) rec[k][x] = 0;
    }
#pragma unroll
    for (int k = 0; k < kFoldRecs; k++) {
      if (r0 + (uint64_t)k * kFoldThreads >= hi) continue;
      const uint32_t g = rec[k][0] & (G - 1), cm = SR::mask(rec[k]);
      s_used[g] = 1;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (!((cm >> a) & 1u)) continue;
        acc_apply_lds(Q::kKind[a], &s_acc[a][g], SR::value(rec[k], a));
        if (Q::kNullable[a]) s_has[null_index<Q>(a)][g] = 1;
      }
    }
  }
  __syncthreads();
  bool ok = true;
  for (uint32_t i = threadIdx.x; i < G && ok; i += kFoldThreads) {
    if (!s_used[i]) continue;
    const int64_t key = (int64_t)((uint64_t)sa.smin + (uint64_t)part * G + i);
    const uint64_t gs = slot_of(sa.t, true, key, &ok);
    if (!ok) break;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      const int64_t v = s_acc[a][i];
      if (v != acc_init(Q::kKind[a]) || Q::kKind[a] == A_SUM_F64) acc_apply_global(Q::kKind[a], &sa.t.acc[a][gs], v);
      if (sa.t.has[a] && (!Q::kNullable[a] || s_has[null_index<Q>(a)][i])) sa.t.has[a][gs] = 1;
    }
  }
  if (!ok) atomicOr(sa.t.overflow, 1u);
}

template <class Q, int S>
__device__ __forceinline__ void aggregate_tiles(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr int LA = Q::kLdsAggs;
  constexpr int LAS = LA > 0 ? LA : 1;
  constexpr int kTile = 256 * S;
  constexpr int HS = Q::kAnyNullable ? kLdsSlots + 2 : 1;  // validity flags only when needed
  constexpr int NLO = Q::kExactF64 ? lo_index<Q>(LA) : 0;  // compensated FP64 sums in LDS
  constexpr bool SP = LA == NA && !Q::kExactF64;            // rows may spill (see AggJitArgs)
  __shared__ uint32_t s_hist[SP ? kMaxSpillParts : 1];
  __shared__ uint32_t s_bucket[SP ? kSpillBuckets : 1];
  __shared__ int64_t s_keys[kLdsSlots + 2];
  __shared__ int64_t s_acc[LAS][kLdsSlots + 2];
  __shared__ int64_t s_lo[NLO > 0 ? NLO : 1][NLO > 0 ? kLdsSlots + 2 : 1];
  __shared__ uint32_t s_has[LAS][HS];
  __shared__ uint32_t s_used;

  const int tid = threadIdx.x, lane = tid & 63;
  for (int i = tid; i < kLdsSlots + 2; i += 256) {
    s_keys[i] = kEmpty;
#pragma unroll
    for (int a = 0; a < LA; a++) {
      s_acc[a][i] = acc_init(Q::kKind[a]);
      if (HS > 1) s_has[a][i] = 0;
    }
#pragma unroll
    for (int b = 0; b < NLO; b++) s_lo[b][i] = 0;
  }
  if (tid == 0) s_used = 0;
  if (SP && p.spill) {
    for (uint32_t i = tid; i < p.sparts; i += 256) s_hist[i] = 0;
    if (tid < kSpillBuckets) s_bucket[tid] = 0;
  }
  __syncthreads();
  uint32_t *const spill_blk = SP ? spill_block<Q>(p) : nullptr;

  bool err = false, ok = true;
  const uint64_t ntiles = (p.n + kTile - 1) / kTile;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    const typename Q::In in_cols = Q::ins(p, base);
    typename Q::Row x[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
    // the block table takes new keys while it is at most half full; after
    // that it only folds the keys it holds (hot keys arrive early and often)
    const bool fill = s_used < kLdsSlots / 2;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      const bool in = j < lim;
      bool sel = in && Q::pred(p.lit, x[s], j, in, err);
      int64_t key = 0;
      bool kv = false;
      int32_t hid_unused;
      if (sel) sel = Q::key(p, x[s], j, in, key, kv, hid_unused);
      if (!kv) key = 0;  // every NULL key is the same group
      int64_t c[NA];
      double cl[NA];  // FP64 sums: the rounding error carried with c
      uint32_t cm = 0;
      Q::contrib(p.lit, x[s], j, sel, c, cm, err);
#pragma unroll
      for (int a = 0; a < NA; a++) cl[a] = 0.0;
      // one matching round: the lanes sharing the first selected lane's key
      // fold into it, so a hot key costs one update per wave
      const uint64_t act = __ballot(sel);
      bool merged = false;  // this lane holds other lanes' rows too
      if (act) {
        const int lead = __builtin_ctzll(act);
        const int64_t lk = (int64_t)__shfl((long long)key, lead, 64);
        const int lkv = __shfl((int)kv, lead, 64);
        const bool same = sel && key == lk && (int)kv == lkv;
        const uint64_t sm = __ballot(same);
        if (__popcll(sm) > 1) {
          uint32_t cmc = 0;
#pragma unroll
          for (int a = 0; a < NA; a++) {
            const bool use = same && ((cm >> a) & 1u);
            int64_t v = use ? c[a] : acc_init(Q::kKind[a]);
            double vl = 0.0;
            switch (Q::kKind[a]) {
              case A_SUM_F64: {
                if (Q::kExactF64) {
                  double hi = as_f64((uint64_t)v);
                  wave_sum_f64(hi, vl);
                  v = (int64_t)f64_bits(hi);
                } else {
                  v = wave_combine<A_SUM_F64>(v);
                }
                break;
              }
              case A_MIN_I64: case A_MIN_F64: case A_MIN_F32: v = wave_combine<A_MIN_I64>(v); break;
              case A_MAX_I64: case A_MAX_F64: case A_MAX_F32: v = wave_combine<A_MAX_I64>(v); break;
              default: v = wave_combine<A_SUM_I64>(v); break;
            }
            if (lane == lead) {  // the other lanes keep their own
              c[a] = v;
              cl[a] = vl;
            }
            cmc |= (uint32_t)(__ballot(use) != 0) << a;
          }
          if (same) {
            if (lane == lead) {
              cm = cmc;
              merged = true;
            } else {
              sel = false;
            }
          }
        }
      }
      // the block table's slot (-1: not held)
      int lslot = -1;
      if (sel && LA > 0) {
        if (!kv) {
          lslot = kLdsSlots;
          s_keys[kLdsSlots] = 0;  // marks the NULL-key slot used
        } else if (key == kEmpty) {
          lslot = kLdsSlots + 1;
          s_keys[kLdsSlots + 1] = 0;
        } else {
          uint32_t h = lds_hash(key);
          for (int pr = 0; pr < kLdsProbe; pr++) {
            const int64_t cur = s_keys[h];
            if (cur == key) {
              lslot = (int)h;
              break;
            }
            if (cur == kEmpty) {
              if (!fill) break;
              const unsigned long long old = atomicCAS((unsigned long long *)&s_keys[h],
                                                       (unsigned long long)kEmpty, (unsigned long long)key);
              if ((int64_t)old == kEmpty) {
                atomicAdd(&s_used, 1u);
                lslot = (int)h;
                break;
              }
              if ((int64_t)old == key) {
                lslot = (int)h;
                break;
              }
            }
            h = (h + 1) & (kLdsSlots - 1);
          }
        }
      }
      // a row the block table does not hold goes to its spill bucket when
      // its key is in the spill range (and the bucket has room), else to the
      // HBM table; a merged row always to HBM (a spill record is one row's:
      // its COUNT is implied and its Int32 values take 32 bits)
      const bool spilled =
          SP && sel && lslot < 0 && !merged && spill_row<Q>(p, spill_blk, s_bucket, s_hist, key, kv, c, cm);
      if (!sel || spilled) continue;
      uint64_t gslot = 0;
      if (lslot < 0 || LA < NA) gslot = slot_of(p.t, kv, key, &ok);
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (!((cm >> a) & 1u)) continue;
        if (a < LA && lslot >= 0) {
          if (Q::kExactF64 && Q::kKind[a] == A_SUM_F64)
            sum_f64(&s_acc[a < LA ? a : 0][lslot], &s_lo[NLO > 0 ? lo_index<Q>(a < LA ? a : 0) : 0][NLO > 0 ? lslot : 0],
                    as_f64((uint64_t)c[a]), cl[a]);
          else
            acc_apply_lds(Q::kKind[a], &s_acc[a < LA ? a : 0][lslot], c[a]);
          if (Q::kNullable[a] && HS > 1) s_has[a < LA ? a : 0][lslot] = 1;
        } else if (ok) {
          if (Q::kExactF64 && Q::kKind[a] == A_SUM_F64)
            sum_f64(&p.t.acc[a][gslot], &p.t.lo[a][gslot], as_f64((uint64_t)c[a]), cl[a]);
          else
            acc_apply_global(Q::kKind[a], &p.t.acc[a][gslot], c[a]);
          if (p.t.has[a]) p.t.has[a][gslot] = 1;
        }
      }
    }
    __syncthreads();  // `fill` of the next tile reads s_used
  }
  // the block table into the HBM table: one update per (group, aggregate)
  if (LA > 0) {
    for (int i = tid; i < kLdsSlots + 2; i += 256) {
      const int64_t k = s_keys[i];
      if (k == kEmpty) continue;
      uint64_t gs;
      if (i == kLdsSlots) gs = slot_of(p.t, false, 0, &ok);
      else if (i == kLdsSlots + 1) gs = slot_of(p.t, true, kEmpty, &ok);
      else gs = global_slot(p.t, k, &ok);
      if (!ok) continue;
#pragma unroll
      for (int a = 0; a < LA; a++) {
        const int64_t v = s_acc[a][i];
        if (Q::kExactF64 && Q::kKind[a] == A_SUM_F64)
          sum_f64(&p.t.acc[a][gs], &p.t.lo[a][gs], as_f64((uint64_t)v),
                  as_f64((uint64_t)s_lo[NLO > 0 ? lo_index<Q>(a) : 0][NLO > 0 ? i : 0]));
        else if (v != acc_init(Q::kKind[a]) || Q::kKind[a] == A_SUM_F64)
          acc_apply_global(Q::kKind[a], &p.t.acc[a][gs], v);
        if (Q::kNullable[a] && HS > 1 && p.t.has[a] && s_has[a][i]) p.t.has[a][gs] = 1;
      }
    }
  }
  if (SP && p.spill) {
    // s_hist and s_bucket are final: every tile ended with a barrier
    for (uint32_t i = tid; i < p.sparts; i += 256) p.hist[(uint64_t)i * gridDim.x + blockIdx.x] = s_hist[i];
    if (tid < kSpillBuckets)
      p.spill_n[blockIdx.x * kSpillBuckets + tid] = s_bucket[tid] < p.spill_cap ? s_bucket[tid] : (uint32_t)p.spill_cap;
  }
  if (err) atomicOr(p.error, 1u);
  if (!ok) atomicOr(p.t.overflow, 1u);
}

// ------------------------------------------------------ hot-group kernel
// For a wide key range with a skewed key distribution (a Zipf fact table
// joined to its dimension): the host samples the input, gives the most
// frequent groups dense hot ids (p.hotmap), and each row of a hot group is
// one LDS atomic per aggregate at its hot id: no hashing, probing or
// matching round. Other rows spill (spill_row) or take the HBM table. Each
// row's join lookup and hot id are loaded for all S row steps before any is
// applied, so those random reads are in flight together. A block's hot
// accumulators end as plain stores into p.hot_part, which a reduce folds
// into the HBM table with one update per (hot group, aggregate). The hot
// map only routes rows: every group's result is the same sum whichever way
// its rows went.
template <class Q, int S>
__device__ __forceinline__ void aggregate_hot(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr int H = hot_capacity(NA, Q::kAnyNullable);
  constexpr int kTile = 256 * S;
  static_assert(NA <= kLdsAggs, "hot groups keep every aggregate in LDS");
  __shared__ int64_t s_acc[NA][H];
  __shared__ uint8_t s_used[H];
  __shared__ uint8_t s_has[Q::kAnyNullable ? NA : 1][Q::kAnyNullable ? H : 1];
  __shared__ uint32_t s_hist[kMaxSpillParts];
  __shared__ uint32_t s_bucket[kSpillBuckets];

  const int tid = threadIdx.x;
  const uint32_t nhot = p.nhot < (uint32_t)H ? p.nhot : (uint32_t)H;  // the host keeps nhot <= H
  for (uint32_t i = tid; i < (uint32_t)H; i += 256) {
    s_used[i] = 0;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      s_acc[a][i] = acc_init(Q::kKind[a]);
      if (Q::kAnyNullable) s_has[a][i] = 0;
    }
  }
  if (p.spill) {
    for (uint32_t i = tid; i < p.sparts; i += 256) s_hist[i] = 0;
    if (tid < kSpillBuckets) s_bucket[tid] = 0;
  }
  __syncthreads();
  uint32_t *const spill_blk = spill_block<Q>(p);

  bool err = false, ok = true;
  const uint64_t ntiles = (p.n + kTile - 1) / kTile;
  auto load_tile = [&](uint64_t tile, typename Q::Row (&x)[S]) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    const typename Q::In in_cols = Q::ins(p, base);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
  };
  // folds tile `tile` from rows x, with the next tile's loads into xn
  // issued first (past the last tile they read `tile` again: unconditional
  // loads keep the compiler's waits on x alone)
  auto step = [&](uint64_t tile, typename Q::Row (&x)[S], typename Q::Row (&xn)[S]) {
    const uint64_t next = tile + gridDim.x;
    load_tile(next < ntiles ? next : tile, xn);
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    // every row step's key (the join lookups, with their hot ids) before any
    // is applied, loads in flight together
    int64_t key[S];
    int32_t hid[S];
    uint32_t selb = 0, kvb = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      const bool in = j < lim;
      bool sel = in && Q::pred(p.lit, x[s], j, in, err);
      int64_t k = 0;
      bool kv = false;
      int32_t h = -1;
      if (sel) sel = Q::key(p, x[s], j, in, k, kv, h);
      if (!kv) k = 0;  // every NULL key is the same group
      key[s] = k;
      hid[s] = h;
      selb |= (uint32_t)sel << s;
      kvb |= (uint32_t)kv << s;
    }
    // keys whose hot id the lookup did not carry: the hot map
#pragma unroll
    for (int s = 0; s < S; s++) {
      if (hid[s] != -2) continue;
      const bool inr = ((kvb >> s) & 1u) && key[s] >= p.smin && (uint64_t)key[s] - (uint64_t)p.smin < p.srange;
      hid[s] = inr ? p.hotmap[(uint64_t)key[s] - (uint64_t)p.smin] : -1;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      const bool sel = (selb >> s) & 1u, kv = (kvb >> s) & 1u;
      int64_t c[NA];
      uint32_t cm = 0;
      Q::contrib(p.lit, x[s], j, sel, c, cm, err);
      if (!sel) continue;
      const int32_t h = hid[s];
      if (h >= 0 && (uint32_t)h < nhot) {
        s_used[h] = 1;
#pragma unroll
        for (int a = 0; a < NA; a++) {
          if (!((cm >> a) & 1u)) continue;
          acc_apply_lds(Q::kKind[a], &s_acc[a][h], c[a]);
          if (Q::kAnyNullable && Q::kNullable[a]) s_has[a][h] = 1;
        }
        continue;
      }
      if (spill_row<Q>(p, spill_blk, s_bucket, s_hist, key[s], kv, c, cm)) continue;
      const uint64_t gslot = slot_of(p.t, kv, key[s], &ok);
      if (!ok) continue;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (!((cm >> a) & 1u)) continue;
        acc_apply_global(Q::kKind[a], &p.t.acc[a][gslot], c[a]);
        if (p.t.has[a]) p.t.has[a][gslot] = 1;
      }
    }
    return next;
  };
  typename Q::Row xa[S], xb[S];
  uint64_t tile = blockIdx.x;
  if (tile < ntiles) {
    load_tile(tile, xa);
    while (true) {
      tile = step(tile, xa, xb);
      if (tile >= ntiles) break;
      tile = step(tile, xb, xa);
      if (tile >= ntiles) break;
    }
  }
  __syncthreads();  // every row applied: the LDS state is final
  // the block's hot partials, for the reduce
  for (uint32_t h = tid; h < nhot; h += 256) {
    uint32_t flags = s_used[h];
#pragma unroll
    for (int a = 0; a < NA; a++) {
      p.hot_part[((uint64_t)blockIdx.x * NA + a) * nhot + h] = s_acc[a][h];
      if (Q::kAnyNullable && Q::kNullable[a]) flags |= (uint32_t)s_has[a][h] << (1 + a);
    }
    p.hot_flags[(uint64_t)blockIdx.x * nhot + h] = (uint8_t)flags;
  }
  if (p.spill) {
    for (uint32_t i = tid; i < p.sparts; i += 256) p.hist[(uint64_t)i * gridDim.x + blockIdx.x] = s_hist[i];
    if (tid < kSpillBuckets)
      p.spill_n[blockIdx.x * kSpillBuckets + tid] = s_bucket[tid] < p.spill_cap ? s_bucket[tid] : (uint32_t)p.spill_cap;
  }
  if (err) atomicOr(p.error, 1u);
  if (!ok) atomicOr(p.t.overflow, 1u);
}

// -------------------------------------------------- dense small-range kernel
// When the key column's value range is small (the host measured it), every
// lane keeps its own accumulator for every (key, aggregate) pair in LDS:
// plain read-modify-writes of the lane's own column, no atomics and no
// probing. The block folds the 256 columns at the end (16 lanes per slot)
// and adds one update per (key, aggregate) to the HBM table. DS slots of 256
// lanes x 8 bytes (DS = 16: 32 KB, DS = 32: 64 KB); in exact mode the FP64
// sums' lo parts take the second half.

__device__ inline int64_t combine_kind(uint8_t kind, int64_t a, int64_t b) {
  switch (kind) {
    case A_SUM_F64: return (int64_t)f64_bits(as_f64((uint64_t)a) + as_f64((uint64_t)b));
    case A_MIN_I64:
    case A_MIN_F64:
    case A_MIN_F32: return a < b ? a : b;
    case A_MAX_I64:
    case A_MAX_F64:
    case A_MAX_F32: return a > b ? a : b;
    default: return (int64_t)((uint64_t)a + (uint64_t)b);
  }
}

// (hi, lo) += (v, vl) with the exact error of hi + v going to lo.
__device__ inline void add_f64_pair(double &hi, double &lo, double v, double vl) {
  const double s = hi + v;
  const double e = __builtin_isfinite(s) ? two_sum_err(hi, v, s) : 0.0;
  hi = s;
  lo = (lo + vl) + e;
}

template <class Q, int S, int DS>
__device__ __forceinline__ void aggregate_dense(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr bool EX = Q::kExactF64;
  constexpr int kTile = 256 * S;
  __shared__ int64_t s_acc[DS * 256];
  __shared__ uint32_t s_bits[2];  // keys used, non-NULL inputs seen

  const int tid = threadIdx.x;
  const uint32_t nkeys = p.dense_keys + p.dense_null;
  const uint32_t nslot = nkeys * NA;  // host: nslot * (EX ? 2 : 1) <= DS, nslot <= 32
  for (uint32_t sl = 0; sl < nslot; sl++) {
    s_acc[sl * 256 + tid] = acc_init(Q::kKind[sl % NA]);
    if (EX) s_acc[(nslot + sl) * 256 + tid] = 0;
  }
  if (tid < 2) s_bits[tid] = 0;
  __syncthreads();  // s_bits before any wave's atomicOr
  uint32_t used = 0, has = 0;
  bool err = false, ok = true;
  const uint64_t ntiles = (p.n + kTile - 1) / kTile;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    const typename Q::In in_cols = Q::ins(p, base);
    typename Q::Row x[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      const bool in = j < lim;
      bool sel = in && Q::pred(p.lit, x[s], j, in, err);
      int64_t key = 0;
      bool kv = false;
      int32_t hid_unused;
      if (sel) sel = Q::key(p, x[s], j, in, key, kv, hid_unused);
      if (!kv) key = 0;
      int64_t c[NA];
      uint32_t cm = 0;
      Q::contrib(p.lit, x[s], j, sel, c, cm, err);
      if (!sel) continue;
      uint32_t r = 0xFFFFFFFFu;
      if (!kv) {
        if (p.dense_null) r = p.dense_keys;
      } else if (key >= p.dense_min && (uint64_t)key - (uint64_t)p.dense_min < p.dense_keys) {
        r = (uint32_t)((uint64_t)key - (uint64_t)p.dense_min);
      }
      if (r == 0xFFFFFFFFu) {  // outside the measured range: the HBM table directly
        const uint64_t gs = slot_of(p.t, kv, key, &ok);
        if (!ok) continue;
#pragma unroll
        for (int a = 0; a < NA; a++) {
          if (!((cm >> a) & 1u)) continue;
          if (EX && Q::kKind[a] == A_SUM_F64) sum_f64(&p.t.acc[a][gs], &p.t.lo[a][gs], as_f64((uint64_t)c[a]), 0.0);
          else acc_apply_global(Q::kKind[a], &p.t.acc[a][gs], c[a]);
          if (p.t.has[a]) p.t.has[a][gs] = 1;
        }
        continue;
      }
      used |= 1u << r;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (!((cm >> a) & 1u)) continue;
        const uint32_t sl = r * NA + a;
        int64_t &acc = s_acc[sl * 256 + tid];
        if (EX && Q::kKind[a] == A_SUM_F64) {
          double hi = as_f64((uint64_t)acc), lo = as_f64((uint64_t)s_acc[(nslot + sl) * 256 + tid]);
          add_f64_pair(hi, lo, as_f64((uint64_t)c[a]), 0.0);
          acc = (int64_t)f64_bits(hi);
          s_acc[(nslot + sl) * 256 + tid] = (int64_t)f64_bits(lo);
        } else {
          acc = combine_kind(Q::kKind[a], acc, c[a]);
        }
        if (Q::kNullable[a]) has |= 1u << sl;
      }
    }
  }
  // the block's key and has bits
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    used |= (uint32_t)__shfl_xor((int)used, d, 64);
    has |= (uint32_t)__shfl_xor((int)has, d, 64);
  }
  if ((tid & 63) == 0) {
    atomicOr(&s_bits[0], used);
    atomicOr(&s_bits[1], has);
  }
  __syncthreads();
  const uint32_t ub = s_bits[0], hb = s_bits[1];
  // fold: 16 lanes per slot, each over 16 of the 256 columns, then across
  // the 16 (xor distances 8..1 stay inside the group)
  for (uint32_t b0 = 0; b0 < nslot; b0 += 16) {
    const uint32_t sl = b0 + (uint32_t)(tid >> 4), part = (uint32_t)(tid & 15);
    const bool live = sl < nslot;
    const uint32_t a = live ? sl % NA : 0, r = live ? sl / NA : 0;
    const uint8_t kind = Q::kKind[a];
    int64_t v = acc_init(kind);
    double vh = 0.0, vl = 0.0;
    if (live) {
      for (uint32_t l = part * 16; l < part * 16 + 16; l++) {
        const int64_t y = s_acc[sl * 256 + l];
        if (EX && kind == A_SUM_F64) add_f64_pair(vh, vl, as_f64((uint64_t)y), as_f64((uint64_t)s_acc[(nslot + sl) * 256 + l]));
        else v = combine_kind(kind, v, y);
      }
    }
#pragma unroll
    for (int d = 8; d > 0; d >>= 1) {
      if (EX && kind == A_SUM_F64) {
        const double oh = __shfl_xor(vh, d, 64), ol = __shfl_xor(vl, d, 64);
        add_f64_pair(vh, vl, oh, ol);
      } else {
        v = combine_kind(kind, v, (int64_t)__shfl_xor((long long)v, d, 64));
      }
    }
    if (live && part == 0 && ((ub >> r) & 1u)) {
      const bool null_slot = p.dense_null && r == p.dense_keys;
      const int64_t key = null_slot ? 0 : (int64_t)((uint64_t)p.dense_min + r);
      const uint64_t gs = slot_of(p.t, !null_slot, key, &ok);
      if (ok) {
        if (EX && kind == A_SUM_F64) sum_f64(&p.t.acc[a][gs], &p.t.lo[a][gs], vh, vl);
        else if (v != acc_init(kind) || kind == A_SUM_F64) acc_apply_global(kind, &p.t.acc[a][gs], v);
        if (p.t.has[a] && ((hb >> sl) & 1u)) p.t.has[a][gs] = 1;
      }
    }
  }
  if (err) atomicOr(p.error, 1u);
  if (!ok) atomicOr(p.t.overflow, 1u);
}


// ------------------------------------------- batched scalar aggregations
// Up to K requests of one prepared shape (the same filter and aggregates
// over one global group, each request with its own literal values) in one
// pass over a table: every tile's rows are loaded once, and each request's
// filter and contributions are evaluated on them into per-lane
// accumulators held in registers. Tiles go to blocks in a fixed order (tile
// += gridDim.x) and every reduction below runs in a fixed order, so a
// request's result depends neither on the requests beside it, nor on the
// batch's size, nor on anything else running on the device; FP64 sums are
// compensated (TwoSum on every row). The run's literals: p.mlits (see
// AggJitArgs). Block b writes request r's partial to
// p.mpart[(b * K + r) * (2 * NA + 1) ...]: NA accumulators, NA lo parts
// (FP64 sums), a flag word (bit a: aggregate a took a value; bit 30: a
// division by zero).

constexpr uint32_t kMultiErr = 1u << 30;

// (acc, lo) of kind `kind` += (v, vl), the compensated add for FP64 sums.
__device__ inline void multi_combine(uint8_t kind, int64_t &acc, double &lo, int64_t v, double vl) {
  if (kind == A_SUM_F64) {
    double hi = as_f64((uint64_t)acc);
    add_f64_pair(hi, lo, as_f64((uint64_t)v), vl);
    acc = (int64_t)f64_bits(hi);
  } else {
    acc = combine_kind(kind, acc, v);
  }
}

// One row's contribution v (when `use`) to a lane's (acc, lo), branch-free:
// an unused row adds the kind's identity. FP64 sums take TwoSum without the
// finiteness guard: once hi is not finite it stays so, and sum_value and the
// pair additions below ignore lo then.
__device__ inline void multi_row(uint8_t kind, int64_t &acc, double &lo, bool use, int64_t v) {
  switch (kind) {
    case A_SUM_F64: {
      const double hi = as_f64((uint64_t)acc);
      const double x = use ? as_f64((uint64_t)v) : 0.0;  // hi is never -0.0: +0.0 is exact
      const double s = hi + x;
      lo += two_sum_err(hi, x, s);
      acc = (int64_t)f64_bits(s);
      break;
    }
    case A_MIN_I64:
    case A_MIN_F64:
    case A_MIN_F32: acc = use && v < acc ? v : acc; break;
    case A_MAX_I64:
    case A_MAX_F64:
    case A_MAX_F32: acc = use && v > acc ? v : acc; break;
    default: acc = (int64_t)((uint64_t)acc + (uint64_t)(use ? v : 0)); break;
  }
}

// Combines (acc, lo, fl) over the wave by a fixed xor tree: every lane ends
// with the same values whatever the order its partners' were made in
// (addition is commutative and TwoSum's error exact).
template <class Q>
__device__ inline void multi_wave(int64_t (&acc)[Q::kAggs], double (&lo)[Q::kAggs], uint32_t &fl) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
#pragma unroll
    for (int a = 0; a < Q::kAggs; a++) {
      const int64_t ov = (int64_t)__shfl_xor((long long)acc[a], d, 64);
      const double ol = Q::kKind[a] == A_SUM_F64 ? __shfl_xor(lo[a], d, 64) : 0.0;
      multi_combine(Q::kKind[a], acc[a], lo[a], ov, ol);
    }
    fl |= (uint32_t)__shfl_xor((int)fl, d, 64);
  }
}

template <class Q, int S, int K>
__device__ __forceinline__ void aggregate_multi(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr int W = 2 * NA + 1;
  constexpr int kTile = 256 * S;
  constexpr int NL = Q::kLits > 0 ? Q::kLits : 1;
  __shared__ int64_t s_acc[4][K][NA];
  __shared__ double s_lo[4][K][NA];
  __shared__ uint32_t s_fl[4][K];
  const int tid = threadIdx.x;
  const uint32_t nreq = (uint32_t)p.mlits[0];
  int64_t acc[K][NA];
  double lo[K][NA];
  uint32_t fl[K];
#pragma unroll
  for (int r = 0; r < K; r++) {
    fl[r] = 0;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      acc[r][a] = acc_init(Q::kKind[a]);
      lo[r][a] = 0.0;
    }
  }
  const uint64_t ntiles = (p.n + kTile - 1) / kTile;
  // the rows of the block's next tile load while this one's are evaluated
  auto load_tile = [&](uint64_t tile, typename Q::Row (&x)[S]) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    const typename Q::In in_cols = Q::ins(p, base);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
  };
  typename Q::Row x[S], xn[S];
  if (blockIdx.x < ntiles) load_tile(blockIdx.x, x);
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    if (tile + gridDim.x < ntiles) load_tile(tile + gridDim.x, xn);
    // contributions that use no literal: once per row for every request,
    // with the row's error (a division by zero) kept apart for the requests
    // that select the row
    // (v0: input a of the row is not NULL, as lane masks the requests'
    // selections AND with)
    int64_t c0[S][NA];
    bool v0[S][NA];
    bool e0[S];
    if constexpr (!Q::kContribLits) {
#pragma unroll
      for (int s = 0; s < S; s++) {
        const uint32_t j = (uint32_t)(s * 256 + tid);
        uint32_t cm0 = 0;
        e0[s] = false;
        Q::contrib(p.mlits, x[s], j, j < lim, c0[s], cm0, e0[s]);
#pragma unroll
        for (int a = 0; a < NA; a++) v0[s][a] = (cm0 >> a) & 1u;
      }
    }
    // every request slot runs, straight-line: the host fills the slots past
    // the run's requests with the last request's literals, and those
    // results are never read
#pragma unroll
    for (int r = 0; r < K; r++) {
      uint64_t L[NL];
#pragma unroll
      for (int i = 0; i < Q::kLits; i++) L[i] = p.mlits[1 + (uint64_t)r * p.mnlit + i];
      bool err = false;
      bool took[NA];
#pragma unroll
      for (int a = 0; a < NA; a++) took[a] = false;
#pragma unroll
      for (int s = 0; s < S; s++) {
        const uint32_t j = (uint32_t)(s * 256 + tid);
        const bool in = j < lim;
        const bool sel = in && Q::pred(L, x[s], j, in, err);
        if constexpr (Q::kContribLits) {
          int64_t c[NA];
          uint32_t cm = 0;
          Q::contrib(L, x[s], j, sel, c, cm, err);
#pragma unroll
          for (int a = 0; a < NA; a++) {
            const bool use = (cm >> a) & 1u;
            multi_row(Q::kKind[a], acc[r][a], lo[r][a], use, c[a]);
            took[a] |= use;
          }
        } else {
#pragma unroll
          for (int a = 0; a < NA; a++) {
            const bool use = sel && v0[s][a];
            multi_row(Q::kKind[a], acc[r][a], lo[r][a], use, c0[s][a]);
            took[a] |= use;
          }
          err |= sel && e0[s];
        }
      }
#pragma unroll
      for (int a = 0; a < NA; a++) fl[r] |= took[a] ? 1u << a : 0u;
      if (err) fl[r] |= kMultiErr;
    }
#pragma unroll
    for (int s = 0; s < S; s++) x[s] = xn[s];
  }
  // the block's partial of each request: waves by a fixed tree, then the
  // four waves in order
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int r = 0; r < K; r++) {
    if ((uint32_t)r < nreq) {
      multi_wave<Q>(acc[r], lo[r], fl[r]);
      if (lane == 0) {
#pragma unroll
        for (int a = 0; a < NA; a++) {
          s_acc[wave][r][a] = acc[r][a];
          s_lo[wave][r][a] = lo[r][a];
        }
        s_fl[wave][r] = fl[r];
      }
    }
  }
  __syncthreads();
  if ((uint32_t)tid < nreq * NA) {
    const int r = tid / NA, a = tid % NA;
    const uint8_t kind = Q::kKind[a];
    int64_t v = s_acc[0][r][a];
    double vl = s_lo[0][r][a];
    for (int w = 1; w < 4; w++) multi_combine(kind, v, vl, s_acc[w][r][a], s_lo[w][r][a]);
    int64_t *o = p.mpart + ((uint64_t)blockIdx.x * K + r) * W;
    o[a] = v;
    o[NA + a] = (int64_t)f64_bits(vl);
  }
  if ((uint32_t)tid < nreq) {
    const uint32_t f = s_fl[0][tid] | s_fl[1][tid] | s_fl[2][tid] | s_fl[3][tid];
    p.mpart[((uint64_t)blockIdx.x * K + tid) * W + 2 * NA] = f;
  }
}

// The final fold: block r (of K) folds request r's partials over the
// p.mblocks blocks in a fixed order and writes its result record to
// p.mres[r * (NA + 1) ...]: each aggregate's value as its output type's
// bits (an FP64 sum as hi + lo, a float MIN/MAX decoded from its order
// key), then the flag word.
template <class Q, int K>
__device__ __forceinline__ void aggregate_multi_final(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr int W = 2 * NA + 1;
  __shared__ int64_t s_acc[4][NA];
  __shared__ double s_lo[4][NA];
  __shared__ uint32_t s_fl[4];
  const uint32_t r = blockIdx.x;
  if (r >= (uint32_t)p.mlits[0]) return;  // uniform over the block
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int64_t acc[NA];
  double lo[NA];
  uint32_t fl = 0;
#pragma unroll
  for (int a = 0; a < NA; a++) {
    acc[a] = acc_init(Q::kKind[a]);
    lo[a] = 0.0;
  }
  for (uint32_t b = (uint32_t)tid; b < p.mblocks; b += 256) {
    const int64_t *src = p.mpart + ((uint64_t)b * K + r) * W;
#pragma unroll
    for (int a = 0; a < NA; a++) multi_combine(Q::kKind[a], acc[a], lo[a], src[a], as_f64((uint64_t)src[NA + a]));
    fl |= (uint32_t)src[2 * NA];
  }
  multi_wave<Q>(acc, lo, fl);
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      s_acc[wave][a] = acc[a];
      s_lo[wave][a] = lo[a];
    }
    s_fl[wave] = fl;
  }
  __syncthreads();
  if (tid < NA) {
    const int a = tid;
    const uint8_t kind = Q::kKind[a];
    int64_t v = s_acc[0][a];
    double vl = s_lo[0][a];
    for (int w = 1; w < 4; w++) multi_combine(kind, v, vl, s_acc[w][a], s_lo[w][a]);
    uint64_t out;
    switch (kind) {
      case A_SUM_F64: out = f64_bits(sum_value(as_f64((uint64_t)v), vl)); break;
      case A_MIN_F64:
      case A_MAX_F64: out = (uint64_t)(v ^ (int64_t)(((uint64_t)(v >> 63)) >> 1)); break;
      case A_MIN_F32:
      case A_MAX_F32: {
        const int32_t k32 = (int32_t)v;
        out = (uint64_t)(uint32_t)(k32 ^ (int32_t)(((uint32_t)(k32 >> 31)) >> 1));
        break;
      }
      default: out = (uint64_t)v;
    }
    p.mres[(uint64_t)r * (NA + 1) + a] = (int64_t)out;
  }
  if (tid == 0) p.mres[(uint64_t)r * (NA + 1) + NA] = s_fl[0] | s_fl[1] | s_fl[2] | s_fl[3];
}

// ------------------------------------------------- keyless aggregation
// A global aggregate (no group key, no join) has one group, so the block
// table and its per-wave matching are pure overhead: each lane keeps its
// own accumulators in registers over every tile it takes (tile +=
// gridDim.x), FP64 sums as (hi, lo) pairs in exact mode; at the end the
// block folds its waves by a fixed tree and then in order, and applies one
// update per aggregate to the HBM table's NULL-key slot (the global group),
// the FP64 sums compensated there as everywhere in exact mode. The slot is
// claimed (seen[0]) only by a block that selected a row, so a global
// aggregate over no selected row still reads COUNT 0 and NULLs.
constexpr uint32_t kKeylessSel = 1u << 31;  // flag bit: the block selected a row

template <class Q, int S>
__device__ __forceinline__ void aggregate_keyless(const AggJitArgs &p) {
  constexpr int NA = Q::kAggs;
  constexpr int kTile = 256 * S;
  __shared__ int64_t s_acc[4][NA];
  __shared__ double s_lo[4][NA];
  __shared__ uint32_t s_fl[4];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int64_t acc[NA];
  double lo[NA];
  uint32_t fl = 0;
#pragma unroll
  for (int a = 0; a < NA; a++) {
    acc[a] = acc_init(Q::kKind[a]);
    lo[a] = 0.0;
  }
  bool err = false;
  const uint64_t ntiles = (p.n + kTile - 1) / kTile;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t base = tile * (uint64_t)kTile;
    const uint32_t lim = (uint32_t)(p.n - base < (uint64_t)kTile ? p.n - base : kTile);
    const typename Q::In in_cols = Q::ins(p, base);
    typename Q::Row x[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * 256 + tid);
      const bool in = j < lim;
      const bool sel = in && Q::pred(p.lit, x[s], j, in, err);
      int64_t c[NA];
      uint32_t cm = 0;
      Q::contrib(p.lit, x[s], j, sel, c, cm, err);
      if (sel) fl |= kKeylessSel;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        if (!((cm >> a) & 1u)) continue;
        if (Q::kKind[a] == A_SUM_F64 && !Q::kExactF64)
          acc[a] = combine_kind(A_SUM_F64, acc[a], c[a]);
        else
          multi_combine(Q::kKind[a], acc[a], lo[a], c[a], 0.0);
        fl |= 1u << a;
      }
    }
  }
  multi_wave<Q>(acc, lo, fl);
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      s_acc[wave][a] = acc[a];
      s_lo[wave][a] = lo[a];
    }
    s_fl[wave] = fl;
  }
  if (err) atomicOr(p.error, 1u);
  __syncthreads();
  const uint32_t f = s_fl[0] | s_fl[1] | s_fl[2] | s_fl[3];
  if (tid < NA && (f & kKeylessSel)) {
    const int a = tid;
    const uint8_t kind = Q::kKind[a];
    int64_t v = s_acc[0][a];
    double vl = s_lo[0][a];
    for (int w = 1; w < 4; w++) multi_combine(kind, v, vl, s_acc[w][a], s_lo[w][a]);
    bool ok = true;
    const uint64_t gs = slot_of(p.t, false, 0, &ok);  // the global group
    if ((f >> a) & 1u) {
      if (Q::kExactF64 && kind == A_SUM_F64)
        sum_f64(&p.t.acc[a][gs], &p.t.lo[a][gs], as_f64((uint64_t)v), vl);
      else if (kind == A_SUM_F64)
        acc_apply_global(kind, &p.t.acc[a][gs], (int64_t)f64_bits(sum_value(as_f64((uint64_t)v), vl)));
      else
        acc_apply_global(kind, &p.t.acc[a][gs], v);
      if (p.t.has[a]) p.t.has[a][gs] = 1;
    }
  }
}

}  // namespace oxr

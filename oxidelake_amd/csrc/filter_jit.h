// Skeleton of the filter + projection kernels that jit.hip generates per
// query shape and compiles at run time with hipRTC. The generated part is a
// struct Q holding the query's straight-line code:
//   struct In, ins(p, base)        the input columns of the tile's rows
//                                  (pointers advanced to its first row, so
//                                  rows are 32-bit offsets j from uniform
//                                  bases: no 64-bit address per lane)
//   struct Row, load(i, j, in, x)  the values (and validity words) one row
//                                  reads (in: the row lies inside its batch)
//   pred(L, x, j, in, err)         the predicate: true keeps the row (L: the
//                                  literals, in prepared mode)
//   struct Out, outs(p, base, at)  the outputs, advanced to the tile's first
//                                  output row `at`; read from the arguments
//                                  once per tile (stores into them could
//                                  otherwise force a reload of every pointer
//                                  at every row step)
//   project(L, o, x, j, q, bitpos, bits, stride, err)
//                                  stores a kept row's projections at output
//                                  row q of the tile and sets its output
//                                  validity bits in LDS
//   vectors(o, sel, j, q, lane)    whole-wave copies of vector projections
//   kValidOuts, kVecOuts           outputs with a validity bitmap / vectors
// Semantics are those of the interpreter kernel in filter.hip
// (crates/oxidelake-device/src/cpu/kernels.rs:99-152); this skeleton does
// the compaction: a tile is B = 64 W threads x S rows, row s*B + tid of the
// tile on thread tid (every load instruction of a wave reads one contiguous run),
// wave ballots rank the kept rows, and a decoupled look-back (compact.h)
// places the tile inside its batch. Compiles under hipcc and hipRTC.
#pragma once

#include "compact.h"

namespace oxr {

constexpr int kJitBlock = 256;
constexpr int kJitCols = 16;
constexpr int kJitOuts = 8;
constexpr int kJitLits = 32;

// Kernel arguments of every generated filter kernel (identical layout on
// the host, which fills it, and in the generated code).
struct JitArgs {
  uint64_t n;                  // input rows
  uint64_t batch_rows;         // rows per batch (each compacted on its own)
  uint64_t tiles_per_batch;
  uint64_t *status;            // one look-back word per tile, zeroed
  uint32_t *counter;           // tile ticket, zeroed
  uint64_t *result;            // [0]: error flags (1 = division by zero)
  uint64_t *counts;            // kept rows per batch
  uint64_t ntiles;             // tiles of the launch (the persistent kernel's ticket bound)
  const void *col[kJitCols];   // input column values
  const uint32_t *valid[kJitCols];
  void *out[kJitOuts];         // output column values
  uint32_t *outv[kJitOuts];    // output validity words (NULL: none)
  int32_t vslot[kJitOuts];     // LDS bit row of output k, -1: none
  int32_t dim[kJitOuts];       // floats per row of a vector output
  int32_t nouts, pad;
  uint64_t lit[kJitLits];      // prepared mode: literal bits, in each literal's type
};

// A column value read once by a generated kernel: nontemporal, so the
// stream does not displace what the caches hold (ld_plain: an ordinary load).
template <typename T>
__device__ inline T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ inline T ld_plain(const T *p) {
  return *p;
}
// A projection value written once (knob jit.stream_stores): nontemporal.
template <typename T>
__device__ inline void st_stream(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// Validity bit of tile row j (tiles start on 32-row boundaries).
__device__ inline bool row_bit(uint32_t w, uint32_t j) { return (w >> (j & 31)) & 1u; }

__device__ inline void set_out_bit(uint32_t *bits, uint32_t bitpos) {
  atomicOr(&bits[bitpos >> 5], 1u << (bitpos & 31));
}

// Copies the dim floats of every kept row of the wave, one row at a time
// with the whole wave; src and dst are advanced to the tile's first input
// and output rows.
__device__ inline void copy_vector_rows(const float *src, float *dst, int dim, bool sel, uint32_t j,
                                        uint32_t q, int lane) {
  uint64_t lanes = __ballot(sel);
  while (lanes) {
    const int l = __builtin_ctzll(lanes);
    lanes &= lanes - 1;
    const uint32_t lj = __shfl(j, l, 64);
    const uint32_t lq = __shfl(q, l, 64);
    const float *s = src + (uint64_t)lj * (uint64_t)dim;
    float *d = dst + (uint64_t)lq * (uint64_t)dim;
    for (int k = lane; k < dim; k += 64) d[k] = s[k];
  }
}

// Rows per lane the skeleton takes: the kept flags of a lane are the bits of
// one 32-bit register.
constexpr int kMaxRowsPerLane = 32;

template <class Q, int S, int W = 4>
__device__ __forceinline__ void filter_tile(const JitArgs &p) {
  static_assert(S >= 1 && S <= kMaxRowsPerLane, "rows per lane: one kept bit per row in a 32-bit register");
  static_assert(W >= 1 && W <= 16, "waves per workgroup");
  constexpr int B = 64 * W;  // threads per workgroup
  // The tile scan runs on one wave over the S * W (step, wave) counts, PL
  // consecutive counts per lane (PL = 1 while S * W <= 64).
  constexpr int NC = S * W;
  constexpr int PL = (NC + 63) / 64;
  constexpr int kTile = B * S;
  constexpr int kWords = kTile / 32 + 2;  // output validity words a tile can touch
  constexpr int kNV = Q::kValidOuts > 0 ? Q::kValidOuts : 1;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_cnt[S * W];
  __shared__ uint32_t s_off[S * W];
  __shared__ uint64_t s_mask[S * W];  // each (step, wave)'s ballot of kept rows
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_bits[kNV * kWords];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(p.counter, 1u);
  if (Q::kValidOuts > 0)
    for (int i = tid; i < kNV * kWords; i += B) s_bits[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;  // ticket order: a batch's tiles in row order
  const uint64_t batch = tile / p.tiles_per_batch;
  const uint64_t lt = tile - batch * p.tiles_per_batch;
  const uint64_t b0 = batch * p.batch_rows;
  const uint64_t bend = b0 + p.batch_rows;
  const uint64_t end = p.n < bend ? p.n : bend;
  const uint64_t base = b0 + lt * (uint64_t)kTile;
  // rows j < lim of the tile exist (a short last batch may leave none)
  const uint32_t lim = end > base ? (uint32_t)(end - base < (uint64_t)kTile ? end - base : kTile) : 0u;

  // 1. every column of every row of the tile in flight at once
  const typename Q::In in_cols = Q::ins(p, base);
  typename Q::Row x[S];
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t j = (uint32_t)(s * B + tid);
    Q::load(in_cols, j, j < lim, x[s]);
  }
  // 2. predicate and ranks. The kept flags live as bits of one register and
  // the ballots in LDS until the stores: as per-lane booleans and 64-bit
  // ballots they would take two SGPR pairs per row step.
  bool err = false;
  uint32_t selbits = 0;
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t j = (uint32_t)(s * B + tid);
    const bool in = j < lim;
    const bool k = Q::pred(p.lit, x[s], j, in, err) && in;
    selbits |= (uint32_t)k << s;
    const uint64_t mk = __ballot(k);
    if (lane == 0) {
      s_cnt[s * W + wave] = (uint32_t)__popcll(mk);
      s_mask[s * W + wave] = mk;
    }
  }
  __syncthreads();
  // 3. tile scan over (step, wave) in row order, then the look-back
  if (wave == 0) {
    uint32_t v[PL], sum = 0;
#pragma unroll
    for (int k = 0; k < PL; k++) {
      const int c = lane * PL + k;
      v[k] = c < NC ? s_cnt[c] : 0u;
      sum += v[k];
    }
    const uint32_t incl = wave_incl_scan(sum, lane);
    uint32_t run = incl - sum;
#pragma unroll
    for (int k = 0; k < PL; k++) {
      const int c = lane * PL + k;
      if (c < NC) s_off[c] = run;
      run += v[k];
    }
    const uint32_t tot = __shfl(incl, 63, 64);
    const uint64_t pre = lookback(p.status, tile, tile - lt, tot, lane);
    if (lane == 0) {
      s_total = tot;
      // a batch's output starts where its input does
      s_prefix = b0 + pre;
      if (lt == p.tiles_per_batch - 1) p.counts[batch] = pre + tot;
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  const uint64_t w0 = prefix >> 5;
  const uint32_t bit0 = (uint32_t)(prefix & 31);
  const uint64_t below = (1ull << lane) - 1;
  // 4. projections of the kept rows, stored at their compacted positions
  const typename Q::Out o = Q::outs(p, base, prefix);
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t j = (uint32_t)(s * B + tid);
    const bool k = (selbits >> s) & 1u;
    const uint32_t q = s_off[s * W + wave] + (uint32_t)__popcll(s_mask[s * W + wave] & below);
    if (k) Q::project(p.lit, o, x[s], j, q, bit0 + q, s_bits, kWords, err);
    if (Q::kVecOuts > 0) Q::vectors(o, k, j, q, lane);
  }
  if (err) atomicOr((unsigned long long *)&p.result[0], 1ull);
  // 5. output validity words: interior words stored, the two boundary
  // words (shared with the neighbouring tiles) OR-ed in
  if (Q::kValidOuts == 0) return;
  const uint32_t tot = s_total;
  if (tot == 0) return;
  __syncthreads();
  const uint64_t last = (prefix + tot - 1) >> 5;
  const int nw = (int)(last - w0 + 1);
  for (int k = 0; k < p.nouts; k++) {
    const int v = p.vslot[k];
    if (v < 0) continue;
    uint32_t *g = p.outv[k];
    for (int w = tid; w < nw; w += B) {
      const uint32_t bits = s_bits[v * kWords + w];
      if (w == 0 || w == nw - 1) {
        if (bits) atomicOr(&g[w0 + w], bits);
      } else {
        g[w0 + w] = bits;
      }
    }
  }
}

// Staged form of filter_tile. The tile's S = C * NCH rows per lane pass
// through registers in NCH chunks of C (the next chunk's loads in flight
// while the current one is ranked), and the kept rows' projections wait in
// LDS at their rank inside the tile rather than in registers: the registers
// hold two chunks instead of the whole tile, so more workgroups share a CU
// and some are loading while others wait on their look-backs. After the
// look-back the staged rows go out as contiguous runs. The stage holds CAP
// rows; from the first chunk whose ranks would pass CAP, the chunks are
// stored directly after the look-back, their rows read a second time (the
// kept flags stay in a register, so the predicate is not evaluated again).
// Output validity bits are collected at their tile rank and shifted to the
// tile's output position when written. Row layout, ranks and results are
// those of filter_tile<Q, C * NCH, W>.
template <class Q, int C, int NCH, int W, int CAP>
__device__ __forceinline__ void filter_staged(const JitArgs &p) {
  constexpr int S = C * NCH;
  static_assert(S >= 1 && S <= kMaxRowsPerLane, "rows per lane: one kept bit per row in a 32-bit register");
  static_assert(W >= 1 && W <= 16, "waves per workgroup");
  static_assert(C * W <= 64, "a chunk's (step, wave) counts: one per lane of one wave");
  static_assert(CAP % 8 == 0, "stage arrays start 8-byte aligned");
  static_assert(Q::kVecOuts == 0, "vector projections are copied by filter_tile");
  constexpr int B = 64 * W;
  constexpr int CC = C * W;
  constexpr int NC = S * W;
  constexpr int kTile = B * S;
  constexpr int kWords = kTile / 32 + 2;
  constexpr int kNV = Q::kValidOuts > 0 ? Q::kValidOuts : 1;
  constexpr int kStage = Q::kStageBytes * CAP > 0 ? Q::kStageBytes * CAP : 8;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_run;     // kept rows of the chunks ranked so far
  __shared__ uint32_t s_staged;  // once the tile passes CAP: ranks below this are in the stage
  __shared__ uint32_t s_cnt[NC];
  __shared__ uint32_t s_off[NC];
  __shared__ uint64_t s_mask[NC];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_bits[kNV * kWords];
  __shared__ __attribute__((aligned(16))) char s_stage[kStage];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(p.counter, 1u);
  if (Q::kValidOuts > 0)
    for (int i = tid; i < kNV * kWords; i += B) s_bits[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t batch = tile / p.tiles_per_batch;
  const uint64_t lt = tile - batch * p.tiles_per_batch;
  const uint64_t b0 = batch * p.batch_rows;
  const uint64_t bend = b0 + p.batch_rows;
  const uint64_t end = p.n < bend ? p.n : bend;
  const uint64_t base = b0 + lt * (uint64_t)kTile;
  const uint32_t lim = end > base ? (uint32_t)(end - base < (uint64_t)kTile ? end - base : kTile) : 0u;
  const uint64_t below = (1ull << lane) - 1;

  const typename Q::In in_cols = Q::ins(p, base);
  const typename Q::Out st = Q::outs_stage(s_stage, CAP);
  typename Q::Row x[C], xn[C];
#pragma unroll
  for (int s = 0; s < C; s++) {
    const uint32_t j = (uint32_t)(s * B + tid);
    Q::load(in_cols, j, j < lim, x[s]);
  }
  bool err = false;
  uint32_t selbits = 0;
  uint32_t run = 0;  // wave 0: kept rows of the chunks ranked so far
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    // predicate and ballots of the chunk's C steps
#pragma unroll
    for (int s = 0; s < C; s++) {
      const int step = ch * C + s;
      const uint32_t j = (uint32_t)(step * B + tid);
      const bool in = j < lim;
      const bool k = Q::pred(p.lit, x[s], j, in, err) && in;
      selbits |= (uint32_t)k << step;
      const uint64_t mk = __ballot(k);
      if (lane == 0) {
        s_cnt[step * W + wave] = (uint32_t)__popcll(mk);
        s_mask[step * W + wave] = mk;
      }
    }
    // the next chunk's loads, in flight over this chunk's ranking and
    // staging (issued after the predicate, so the loads a wait covers are
    // one chunk's, within what the load counter tracks)
    if (ch + 1 < NCH) {
#pragma unroll
      for (int s = 0; s < C; s++) {
        const uint32_t j = (uint32_t)(((ch + 1) * C + s) * B + tid);
        Q::load(in_cols, j, j < lim, xn[s]);
      }
    }
    __syncthreads();
    // the chunk's ranks, continuing the tile's running count
    if (wave == 0) {
      const uint32_t v = lane < CC ? s_cnt[ch * CC + lane] : 0u;
      const uint32_t incl = wave_incl_scan(v, lane);
      if (lane < CC) s_off[ch * CC + lane] = run + incl - v;
      const uint32_t prev = run;
      run += __shfl(incl, 63, 64);
      if (lane == 0) {
        s_run = run;
        // the first chunk past the stage starts where the stage ends
        if (prev <= (uint32_t)CAP && run > (uint32_t)CAP) s_staged = prev;
      }
    }
    __syncthreads();
    // a chunk goes to the stage when all its ranks fit (ranks only grow,
    // so the staged chunks are a prefix of the tile)
    if (s_run <= (uint32_t)CAP) {
#pragma unroll
      for (int s = 0; s < C; s++) {
        const int step = ch * C + s;
        if ((selbits >> step) & 1u) {
          const uint32_t j = (uint32_t)(step * B + tid);
          const uint32_t q = s_off[step * W + wave] + (uint32_t)__popcll(s_mask[step * W + wave] & below);
          Q::project(p.lit, st, x[s], j, q, q, s_bits, kWords, err);
        }
      }
    }
    if (ch + 1 < NCH) {
#pragma unroll
      for (int s = 0; s < C; s++) x[s] = xn[s];
    }
  }
  // look-back over the tile's total
  if (wave == 0) {
    const uint64_t pre = lookback(p.status, tile, tile - lt, run, lane);
    if (lane == 0) {
      s_prefix = b0 + pre;
      if (lt == p.tiles_per_batch - 1) p.counts[batch] = pre + run;
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  const uint32_t tot = s_run;
  const uint32_t staged = tot <= (uint32_t)CAP ? tot : s_staged;
  const typename Q::Out o = Q::outs(p, base, prefix);
  // staged rows out as contiguous runs
  Q::copy_out(o, s_stage, CAP, staged, tid, B);
  // chunks past the stage: rows read again, kept rows stored directly
  if (staged < tot) {
    for (int ch = 0; ch < NCH; ch++) {
      if (s_off[ch * CC] < staged) continue;
#pragma unroll
      for (int s = 0; s < C; s++) {
        const uint32_t j = (uint32_t)((ch * C + s) * B + tid);
        Q::load(in_cols, j, j < lim, x[s]);
      }
#pragma unroll
      for (int s = 0; s < C; s++) {
        const int step = ch * C + s;
        if ((selbits >> step) & 1u) {
          const uint32_t j = (uint32_t)(step * B + tid);
          const uint32_t q = s_off[step * W + wave] + (uint32_t)__popcll(s_mask[step * W + wave] & below);
          Q::project(p.lit, o, x[s], j, q, q, s_bits, kWords, err);
        }
      }
    }
  }
  if (err) atomicOr((unsigned long long *)&p.result[0], 1ull);
  // output validity words: bit q of the tile is output bit prefix + q
  if (Q::kValidOuts == 0 || tot == 0) return;
  __syncthreads();
  const uint64_t w0 = prefix >> 5;
  const uint32_t bit0 = (uint32_t)(prefix & 31);
  const int nw = (int)(((prefix + tot - 1) >> 5) - w0 + 1);
  for (int k = 0; k < p.nouts; k++) {
    const int v = p.vslot[k];
    if (v < 0) continue;
    uint32_t *g = p.outv[k];
    const uint32_t *t = s_bits + v * kWords;
    for (int w = tid; w < nw; w += B) {
      const uint32_t hi = t[w];
      const uint32_t bits = bit0 == 0 ? hi : (hi << bit0) | (w > 0 ? t[w - 1] >> (32 - bit0) : 0u);
      if (w == 0 || w == nw - 1) {
        if (bits) atomicOr(&g[w0 + w], bits);
      } else {
        g[w0 + w] = bits;
      }
    }
  }
}

// Persistent form of filter_tile: a grid of about as many workgroups as fit
// on the device at once, each taking tiles by ticket until none is left.
// Two register sets of S rows: the next tile's loads are issued before the
// current tile's predicate, look-back and stores, so a workgroup keeps
// loads in flight while it compacts (one tile per workgroup launch leaves
// the loads and the compaction of the resident workgroups in step). A
// workgroup takes the ticket of its next tile when it starts the current
// one, so the tiles compacted at any time are one round of tickets and a
// tile's look-back mostly waits on tiles compacted beside it (tickets taken
// two at a time up front chained every tile to a tile its owner had not
// started: 3x slower). A look-back waits only on tiles with smaller
// tickets, each held by a running workgroup that reaches it without waiting
// on a later tile, so every wait ends.
template <class Q, int S, int W = 4>
__device__ __forceinline__ void filter_persistent(const JitArgs &p) {
  static_assert(S >= 1 && S <= kMaxRowsPerLane, "rows per lane: one kept bit per row in a 32-bit register");
  static_assert(W >= 1 && W <= 16, "waves per workgroup");
  constexpr int B = 64 * W;  // threads per workgroup
  constexpr int NC = S * W;
  constexpr int PL = (NC + 63) / 64;
  constexpr int kTile = B * S;
  constexpr int kWords = kTile / 32 + 2;
  constexpr int kNV = Q::kValidOuts > 0 ? Q::kValidOuts : 1;
  __shared__ uint32_t s_tile, s_next;
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_cnt[NC];
  __shared__ uint32_t s_off[NC];
  __shared__ uint64_t s_mask[NC];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_bits[kNV * kWords];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(p.counter, 1u);
  if (Q::kValidOuts > 0)
    for (int i = tid; i < kNV * kWords; i += B) s_bits[i] = 0;
  __syncthreads();
  uint64_t cur = s_tile;
  if (cur >= p.ntiles) return;  // uniform: no tile for this workgroup

  // the rows [base, base + lim) of tile t
  auto tile_rows = [&](uint64_t t, uint64_t &base, uint32_t &lim) {
    const uint64_t batch = t / p.tiles_per_batch;
    const uint64_t lt = t - batch * p.tiles_per_batch;
    const uint64_t b0 = batch * p.batch_rows;
    const uint64_t bend = b0 + p.batch_rows;
    const uint64_t end = p.n < bend ? p.n : bend;
    base = b0 + lt * (uint64_t)kTile;
    lim = end > base ? (uint32_t)(end - base < (uint64_t)kTile ? end - base : kTile) : 0u;
  };
  auto load_tile = [&](uint64_t t, typename Q::Row (&x)[S]) {
    uint64_t base;
    uint32_t lim;
    tile_rows(t, base, lim);
    const typename Q::In in_cols = Q::ins(p, base);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
  };
  // takes the next tile's ticket, issues its loads into xn, compacts tile t
  // from rows x; returns the next tile. Past the last tile the loads read
  // tile t again (in L2 by now): unconditional loads let the compiler wait
  // on x's loads alone while xn's stay in flight.
  bool err = false;
  auto step = [&](uint64_t t, typename Q::Row (&x)[S], typename Q::Row (&xn)[S]) {
    if (tid == 0) s_next = atomicAdd(p.counter, 1u);
    __syncthreads();  // the ticket visible; the previous tile's reads of LDS done
    const uint64_t next = s_next;
    load_tile(next < p.ntiles ? next : t, xn);
    uint64_t base;
    uint32_t lim;
    tile_rows(t, base, lim);
    const uint64_t batch = t / p.tiles_per_batch;
    const uint64_t lt = t - batch * p.tiles_per_batch;
    const uint64_t b0 = batch * p.batch_rows;
    uint32_t selbits = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      const bool in = j < lim;
      const bool k = Q::pred(p.lit, x[s], j, in, err) && in;
      selbits |= (uint32_t)k << s;
      const uint64_t mk = __ballot(k);
      if (lane == 0) {
        s_cnt[s * W + wave] = (uint32_t)__popcll(mk);
        s_mask[s * W + wave] = mk;
      }
    }
    __syncthreads();  // ballots in LDS
    if (wave == 0) {
      uint32_t v[PL], sum = 0;
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int c = lane * PL + k;
        v[k] = c < NC ? s_cnt[c] : 0u;
        sum += v[k];
      }
      const uint32_t incl = wave_incl_scan(sum, lane);
      uint32_t run = incl - sum;
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int c = lane * PL + k;
        if (c < NC) s_off[c] = run;
        run += v[k];
      }
      const uint32_t tot = __shfl(incl, 63, 64);
      const uint64_t pre = lookback(p.status, t, t - lt, tot, lane);
      if (lane == 0) {
        s_total = tot;
        s_prefix = b0 + pre;
        if (lt == p.tiles_per_batch - 1) p.counts[batch] = pre + tot;
      }
    }
    __syncthreads();  // offsets and prefix visible
    const uint64_t prefix = s_prefix;
    const uint64_t w0 = prefix >> 5;
    const uint32_t bit0 = (uint32_t)(prefix & 31);
    const uint64_t below = (1ull << lane) - 1;
    const typename Q::Out o = Q::outs(p, base, prefix);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      const bool k = (selbits >> s) & 1u;
      const uint32_t q = s_off[s * W + wave] + (uint32_t)__popcll(s_mask[s * W + wave] & below);
      if (k) Q::project(p.lit, o, x[s], j, q, bit0 + q, s_bits, kWords, err);
      if (Q::kVecOuts > 0) Q::vectors(o, k, j, q, lane);
    }
    const uint32_t tot = s_total;
    __syncthreads();  // every output bit set; every read of s_off, s_mask, s_total done
    if (Q::kValidOuts > 0 && tot > 0) {
      const uint64_t last = (prefix + tot - 1) >> 5;
      const int nw = (int)(last - w0 + 1);
      for (int k = 0; k < p.nouts; k++) {
        const int v = p.vslot[k];
        if (v < 0) continue;
        uint32_t *g = p.outv[k];
        for (int w = tid; w < nw; w += B) {
          const uint32_t bits = s_bits[v * kWords + w];
          s_bits[v * kWords + w] = 0;  // clear for the next tile (read only by this thread)
          if (w == 0 || w == nw - 1) {
            if (bits) atomicOr(&g[w0 + w], bits);
          } else {
            g[w0 + w] = bits;
          }
        }
      }
    }
    return next;
  };
  typename Q::Row xa[S], xb[S];
  load_tile(cur, xa);
  while (true) {
    cur = step(cur, xa, xb);
    if (cur >= p.ntiles) break;
    cur = step(cur, xb, xa);
    if (cur >= p.ntiles) break;
  }
  if (err) atomicOr((unsigned long long *)&p.result[0], 1ull);
}


// One workgroup per batch (many-batch inputs): the workgroups take batches
// blockIdx.x, blockIdx.x + gridDim.x, ... and walk each batch's tiles in row
// order, so a tile's place in its batch is the running count of the tiles
// before it: no look-back and no ticket. Two register sets of S rows: the
// next tile's loads are issued before the current tile is ranked, projected
// and stored, so the loads of one workgroup run under its own compaction.
// Needs as many batches as there are workgroups to fill the device with;
// the host picks it for inputs of many batches (knob filter.wg).
template <class Q, int S, int W = 4>
__device__ __forceinline__ void filter_batch_wg(const JitArgs &p) {
  static_assert(S >= 1 && S <= kMaxRowsPerLane, "rows per lane: one kept bit per row in a 32-bit register");
  static_assert(W >= 1 && W <= 16, "waves per workgroup");
  constexpr int B = 64 * W;  // threads per workgroup
  constexpr int NC = S * W;
  constexpr int PL = (NC + 63) / 64;
  constexpr int kTile = B * S;
  constexpr int kWords = kTile / 32 + 2;
  constexpr int kNV = Q::kValidOuts > 0 ? Q::kValidOuts : 1;
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_cnt[NC];
  __shared__ uint32_t s_off[NC];
  __shared__ uint64_t s_mask[NC];
  __shared__ uint32_t s_bits[kNV * kWords];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint64_t nb = (p.n + p.batch_rows - 1) / p.batch_rows;
  const uint64_t T = p.tiles_per_batch;
  uint64_t b = blockIdx.x;
  if (b >= nb) return;  // uniform
  if (Q::kValidOuts > 0)
    for (int i = tid; i < kNV * kWords; i += B) s_bits[i] = 0;

  // the rows [base, base + lim) of tile lt of batch bb
  auto tile_rows = [&](uint64_t bb, uint64_t lt, uint64_t &base, uint32_t &lim) {
    const uint64_t b0 = bb * p.batch_rows;
    const uint64_t bend = b0 + p.batch_rows;
    const uint64_t end = p.n < bend ? p.n : bend;
    base = b0 + lt * (uint64_t)kTile;
    lim = end > base ? (uint32_t)(end - base < (uint64_t)kTile ? end - base : kTile) : 0u;
  };
  auto load_tile = [&](uint64_t bb, uint64_t lt, typename Q::Row (&x)[S]) {
    uint64_t base;
    uint32_t lim;
    tile_rows(bb, lt, base, lim);
    const typename Q::In in_cols = Q::ins(p, base);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      Q::load(in_cols, j, j < lim, x[s]);
    }
  };
  bool err = false;
  uint64_t run = 0;  // kept rows of the batch's tiles before the current one
  uint64_t lt = 0;
  // compacts tile lt of batch b from rows x, the next tile's loads going
  // into xn (past the workgroup's last tile they read the current tile
  // again, from L2: unconditional loads let the compiler wait on x's loads
  // alone while xn's stay in flight); advances (b, lt, run)
  auto step = [&](typename Q::Row (&x)[S], typename Q::Row (&xn)[S]) {
    const bool last_tile = lt + 1 == T;
    const uint64_t nbb = last_tile ? b + gridDim.x : b;
    const uint64_t nlt = last_tile ? 0 : lt + 1;
    if (nbb < nb) load_tile(nbb, nlt, xn);
    else load_tile(b, lt, xn);
    uint64_t base;
    uint32_t lim;
    tile_rows(b, lt, base, lim);
    const uint64_t b0 = b * p.batch_rows;
    uint32_t selbits = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      const bool in = j < lim;
      const bool k = Q::pred(p.lit, x[s], j, in, err) && in;
      selbits |= (uint32_t)k << s;
      const uint64_t mk = __ballot(k);
      if (lane == 0) {
        s_cnt[s * W + wave] = (uint32_t)__popcll(mk);
        s_mask[s * W + wave] = mk;
      }
    }
    __syncthreads();  // ballots in LDS
    if (wave == 0) {
      uint32_t v[PL], sum = 0;
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int c = lane * PL + k;
        v[k] = c < NC ? s_cnt[c] : 0u;
        sum += v[k];
      }
      const uint32_t incl = wave_incl_scan(sum, lane);
      uint32_t r = incl - sum;
#pragma unroll
      for (int k = 0; k < PL; k++) {
        const int c = lane * PL + k;
        if (c < NC) s_off[c] = r;
        r += v[k];
      }
      const uint32_t tot = __shfl(incl, 63, 64);  // the whole wave: lane 63's total
      if (lane == 0) s_total = tot;
    }
    __syncthreads();  // offsets and total visible
    const uint64_t prefix = b0 + run;  // a batch's output starts where its input does
    const uint64_t w0 = prefix >> 5;
    const uint32_t bit0 = (uint32_t)(prefix & 31);
    const uint64_t below = (1ull << lane) - 1;
    const typename Q::Out o = Q::outs(p, base, prefix);
#pragma unroll
    for (int s = 0; s < S; s++) {
      const uint32_t j = (uint32_t)(s * B + tid);
      const bool k = (selbits >> s) & 1u;
      const uint32_t q = s_off[s * W + wave] + (uint32_t)__popcll(s_mask[s * W + wave] & below);
      if (k) Q::project(p.lit, o, x[s], j, q, bit0 + q, s_bits, kWords, err);
      if (Q::kVecOuts > 0) Q::vectors(o, k, j, q, lane);
    }
    const uint32_t tot = s_total;
    __syncthreads();  // every output bit set; every read of s_off, s_mask, s_total done
    if (Q::kValidOuts > 0 && tot > 0) {
      const uint64_t last = (prefix + tot - 1) >> 5;
      const int nw = (int)(last - w0 + 1);
      for (int k = 0; k < p.nouts; k++) {
        const int v = p.vslot[k];
        if (v < 0) continue;
        uint32_t *g = p.outv[k];
        for (int w = tid; w < nw; w += B) {
          const uint32_t bits = s_bits[v * kWords + w];
          s_bits[v * kWords + w] = 0;  // clear for the next tile (read only by this thread)
          if (w == 0 || w == nw - 1) {
            if (bits) atomicOr(&g[w0 + w], bits);
          } else {
            g[w0 + w] = bits;
          }
        }
      }
    }
    run += tot;
    if (last_tile) {
      if (tid == 0) p.counts[b] = run;
      run = 0;
    }
    b = nbb;
    lt = nlt;
  };
  typename Q::Row xa[S], xb[S];
  load_tile(b, 0, xa);
  __syncthreads();  // s_bits cleared
  while (true) {
    step(xa, xb);
    if (b >= nb) break;
    step(xb, xa);
    if (b >= nb) break;
  }
  if (err) atomicOr((unsigned long long *)&p.result[0], 1ull);
}

}  // namespace oxr

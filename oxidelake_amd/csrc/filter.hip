// Fused filter + projection with stream compaction, in one pass.
//
// Replaces `ops::filter_project` (crates/oxidelake-device/src/cuda/ops.rs:
// 143-373), which launches one compare kernel per predicate leaf, one mask
// AND per conjunction, a block count, a device->host copy of the counts with
// a host scan, a scatter of row indices and one gather per projected column.
// Here one kernel per batch:
//   1. each lane loads every column the query reads for its 2 x 4
//      consecutive rows with 16-byte vector loads, all in flight at once;
//   2. it evaluates the whole predicate in registers (expr.h);
//   3. a wave64 prefix sum of the per-lane counts plus a block scan give each
//      selected row its rank inside the tile;
//   4. a single-pass decoupled look-back over the tiles (tile ids taken from
//      an atomic counter, so every tile a lane waits on is already running)
//      gives the tile's output offset — no host round trip;
//   5. each lane computes the projections of its selected rows from the same
//      registers and stores them at their compacted positions; output
//      validity bits are assembled in LDS and written as whole words
//      (boundary words with atomicOr).
// Semantics: crates/oxidelake-device/src/cpu/kernels.rs:99-152.
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>

#include "common.h"
#include "compact.h"
#include "expr.h"
#include "jit.h"

namespace oxr {
namespace {

// Launch shape: V consecutive rows per lane, G row groups per lane, so a
// tile is kBlock * V * G rows. Shapes are instantiated below and picked per
// call (the autotuner's choice, else the default).
constexpr int kBlock = 256;      // threads per block (4 waves)
constexpr int kWaves = kBlock / 64;

struct FilterShape {
  int v, g;
};

// The instantiated shapes; knobs "filter.v"/"filter.g" pick one (default
// one row per lane, two row groups: at most ~90 VGPRs for four columns). A
// pair that is not instantiated is refused, never replaced.
int filter_shape(FilterShape *out) {
  const int v = (int)tune_get("filter.v", 1), g = (int)tune_get("filter.g", 2);
  const FilterShape known[] = {{1, 1}, {1, 2}, {1, 4}, {2, 1}, {2, 2}, {4, 1}};
  for (const FilterShape &k : known)
    if (k.v == v && k.g == g) {
      *out = k;
      return OXR_OK;
    }
  return fail(OXR_EPLAN, "rocm.tune",
              "filter.v x filter.g = " + std::to_string(v) + " x " + std::to_string(g) +
                  " is not an instantiated shape (v * g <= 4)");
}

struct Out {
  void *values;
  uint32_t *valid;  // NULL when the projection cannot be NULL
  ValueProg vp;     // slot, or program; vp.col for vector columns
  int32_t dim;
};

struct Params {
  uint64_t n;
  uint64_t ntiles;
  DCol cols[kMaxCols];
  Ins ins[kMaxIns];
  Leaf leaves[kMaxLeaves];
  uint8_t bprog[kMaxBool];
  DivMagic magic[kMaxMagic];
  int8_t slot_cols[kMaxSlots];
  int32_t nslots;
  Out outs[kMaxOuts];
  int32_t nb, nouts;
  uint64_t *status;   // ntiles words, zeroed
  uint32_t *counter;  // tile ticket, zeroed
  uint64_t *result;   // [0] = error flags
  uint64_t *counts;   // selected rows per batch
  uint64_t batch_rows, tiles_per_batch;
};

__device__ inline void store_value(void *base, int type, uint64_t pos, uint64_t v) {
  if (type == OXR_INT32 || type == OXR_FLOAT32) ((uint32_t *)base)[pos] = (uint32_t)v;
  else ((uint64_t *)base)[pos] = v;
}

template <int NS, int V, int kGroups>
__global__ __launch_bounds__(kBlock) void filter_project_kernel(const Params p) {
  constexpr int kTile = kBlock * V * kGroups;
  constexpr int kTileWords = kTile / 32 + 2;  // output validity words a tile can touch
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wtot[kGroups][kWaves];
  __shared__ uint32_t s_woff[kGroups][kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_bits[kMaxOuts][kTileWords];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(p.counter, 1u);
  for (int i = tid; i < kMaxOuts * kTileWords; i += kBlock) (&s_bits[0][0])[i] = 0;
  __syncthreads();
  const uint64_t tile = s_tile;  // ticket order: a batch's tiles in row order
  const uint64_t batch = tile / p.tiles_per_batch;
  const uint64_t lt = tile - batch * p.tiles_per_batch;
  const uint64_t batch_row0 = batch * p.batch_rows;
  const uint64_t row_end = min(p.n, batch_row0 + p.batch_rows);
  const uint64_t tile_base = batch_row0 + lt * (uint64_t)kTile;
  bool div0 = false;

  // 1. every column of every row group in flight at once
  Slots<V, NS> S[kGroups];
#pragma unroll
  for (int j = 0; j < kGroups; j++) {
    const uint64_t row0 = tile_base + ((uint64_t)(j * kBlock + tid)) * V;
    load_slots<V, NS>(p.cols, p.slot_cols, p.nslots, row0, row_end, S[j]);
  }
  // 2. predicate
  uint32_t sel[kGroups];
  uint32_t excl[kGroups];
#pragma unroll
  for (int j = 0; j < kGroups; j++) {
    const uint64_t row0 = tile_base + ((uint64_t)(j * kBlock + tid)) * V;
    sel[j] = eval_predicate<V, NS>(p.leaves, p.bprog, p.nb, p.ins, p.cols, S[j], p.magic, row0,
                                   row_end, &div0);
    if constexpr (V == 1) {
      // one row per lane: ranks straight from the wave's ballot
      const uint64_t b = __ballot(sel[j] != 0);
      excl[j] = (uint32_t)__popcll(b & ((1ull << lane) - 1));
      if (lane == 0) s_wtot[j][wave] = (uint32_t)__popcll(b);
    } else {
      const uint32_t cnt = __builtin_popcount(sel[j]);
      const uint32_t incl = wave_incl_scan(cnt, lane);
      excl[j] = incl - cnt;
      if (lane == 63) s_wtot[j][wave] = incl;
    }
  }
  __syncthreads();
  // 3. block scan over (group, wave) in row order
  if (tid == 0) {
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kGroups; j++)
#pragma unroll
      for (int w = 0; w < kWaves; w++) {
        s_woff[j][w] = run;
        run += s_wtot[j][w];
      }
    s_total = run;
  }
  __syncthreads();
  // 4. look-back
  if (wave == 0) {
    const uint64_t pre = lookback(p.status, tile, tile - lt, s_total, lane);
    if (lane == 0) {
      // a batch's output starts where its input does: batches need no
      // offsets from one another
      s_prefix = batch_row0 + pre;
      if (lt == p.tiles_per_batch - 1) p.counts[batch] = pre + s_total;
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  const uint64_t word0 = prefix >> 5;

  // 5. projections, of the kept rows only: a lane's other rows must not
  // raise errors (a zero divisor in a dropped row is no error)
#pragma unroll
  for (int j = 0; j < kGroups; j++)
#pragma unroll
    for (int k = 0; k < NS; k++) S[j].m[k] &= sel[j];
  for (int o = 0; o < p.nouts; o++) {
    const Out out = p.outs[o];
    const int type = out.vp.type;
#pragma unroll
    for (int j = 0; j < kGroups; j++) {
      const uint64_t row0 = tile_base + ((uint64_t)(j * kBlock + tid)) * V;
      uint64_t pos = prefix + s_woff[j][wave] + excl[j];
      if (type == OXR_F32_VECTOR) {
        // whole-wave copy of each selected row's dim floats
        const int dim = out.dim;
        const DCol src_col = p.cols[out.vp.col];
        const float *src = (const float *)src_col.values;
        float *dst = (float *)out.values;
        uint64_t lanes = __ballot(sel[j] != 0);
        while (lanes) {
          const int l = __builtin_ctzll(lanes);
          lanes &= lanes - 1;
          const uint32_t ls = __shfl(sel[j], l, 64);
          uint64_t lpos = __shfl(pos, l, 64);
          const uint64_t lrow0 = __shfl(row0, l, 64);
          for (int q = 0; q < V; q++) {
            if (!((ls >> q) & 1)) continue;
            const float *s = src + (lrow0 + q) * (uint64_t)dim;
            float *d = dst + lpos * (uint64_t)dim;
            for (int k = lane; k < dim; k += 64) d[k] = s[k];
            lpos++;
          }
        }
        if (out.valid && src_col.valid) {
          uint64_t at = pos;
          for (int q = 0; q < V; q++) {
            if (!((sel[j] >> q) & 1)) continue;
            const uint64_t r = row0 + q;
            if ((src_col.valid[r >> 5] >> (r & 31)) & 1)
              atomicOr(&s_bits[o][(at >> 5) - word0], 1u << (at & 31));
            at++;
          }
        }
        continue;
      }
      if (!sel[j]) continue;
      uint64_t v[V];
      uint32_t vm = 0;
      eval_value<V, NS>(out.vp, p.ins, S[j], p.magic, v, vm, &div0);
#pragma unroll
      for (int q = 0; q < V; q++) {
        if (!((sel[j] >> q) & 1)) continue;
        store_value(out.values, type, pos, v[q]);
        if (out.valid && ((vm >> q) & 1)) atomicOr(&s_bits[o][(pos >> 5) - word0], 1u << (pos & 31));
        pos++;
      }
    }
  }
  if (div0) atomicOr((unsigned long long *)&p.result[0], 1ull);
  // 6. validity words
  const uint32_t total = s_total;
  bool any_valid = false;
  for (int o = 0; o < p.nouts; o++) any_valid |= p.outs[o].valid != nullptr;
  if (!any_valid || total == 0) return;
  __syncthreads();
  const uint64_t last = (prefix + total - 1) >> 5;
  const int nwords = (int)(last - word0 + 1);
  for (int o = 0; o < p.nouts; o++) {
    uint32_t *g = p.outs[o].valid;
    if (!g) continue;
    for (int w = tid; w < nwords; w += kBlock) {
      const uint32_t bits = s_bits[o][w];
      if (w == 0 || w == nwords - 1) {
        if (bits) atomicOr(&g[word0 + w], bits);
      } else {
        g[word0 + w] = bits;
      }
    }
  }
}

template <int NS>
void launch_ns(const FilterShape &sh, dim3 grid, hipStream_t st, const Params &P) {
  const dim3 block(kBlock);
  if (sh.v == 1 && sh.g == 1) hipLaunchKernelGGL((filter_project_kernel<NS, 1, 1>), grid, block, 0, st, P);
  else if (sh.v == 1 && sh.g == 4) hipLaunchKernelGGL((filter_project_kernel<NS, 1, 4>), grid, block, 0, st, P);
  else if (sh.v == 2 && sh.g == 1) hipLaunchKernelGGL((filter_project_kernel<NS, 2, 1>), grid, block, 0, st, P);
  else if (sh.v == 2 && sh.g == 2) hipLaunchKernelGGL((filter_project_kernel<NS, 2, 2>), grid, block, 0, st, P);
  else if (sh.v == 4 && sh.g == 1) hipLaunchKernelGGL((filter_project_kernel<NS, 4, 1>), grid, block, 0, st, P);
  else hipLaunchKernelGGL((filter_project_kernel<NS, 1, 2>), grid, block, 0, st, P);
}

void launch_filter(int ns, const FilterShape &sh, dim3 grid, hipStream_t st, const Params &P) {
  switch (ns) {
    case 2: launch_ns<2>(sh, grid, st, P); break;
    case 4: launch_ns<4>(sh, grid, st, P); break;
    default: launch_ns<8>(sh, grid, st, P); break;
  }
}

}  // namespace
}  // namespace oxr

using namespace oxr;

namespace oxr {
namespace {

// Where a filter's outputs and scratch come from: by default the arena;
// for the no-allocation entry point the caller's output columns and the
// static scratch array, and only kernels already loaded.
struct FilterTarget {
  const oxr_column *outs = nullptr;  // nproj caller-owned output columns
  void *scratch = nullptr;           // device scratch of scratch_bytes
  size_t scratch_bytes = 0;
};

// Runs the filter over `input` cut into batches of `batch_rows` rows, all in
// one launch. Output column buffers have the input's capacity; batch i's
// output rows start at row i * batch_rows of them. counts[i] = its rows.
int run_filter(oxr_stream *s, const oxr_batch *input, uint64_t batch_rows, const oxr_expr *predicate,
               const oxr_expr *projection, int32_t nproj, const char *const *names,
               std::vector<DevCol> *out_cols, std::vector<std::string> *out_names,
               std::vector<uint64_t> *counts, const FilterTarget *tgt = nullptr) {
  if (nproj < 0 || nproj > kMaxOuts)
    return fail(OXR_EUNSUPPORTED, "rocm.filter",
                "at most " + std::to_string(kMaxOuts) + " projections per launch");
  if ((int)input->cols.size() > kMaxCols)
    return fail(OXR_EUNSUPPORTED, "rocm.filter",
                "at most " + std::to_string(kMaxCols) + " input columns per launch");
  Program prog;
  OXR_TRY(compile_predicate(predicate, input->cols, &prog));
  std::vector<ValueProg> vps(nproj);
  for (int i = 0; i < nproj; i++) OXR_TRY(compile_value(&projection[i], input->cols, &prog, &vps[i]));

  // the query's own generated kernel, else the interpreter
  JitFilter jf;
  bool use_jit = false;
  if (tune_get("filter.jit", 1) != 0 && input->rows > 0) {
    const int jrc = jit_filter(s->dev, input->cols, predicate, projection, nproj, vps, &jf, tgt != nullptr);
    if (jrc == OXR_OK) use_jit = true;
    else if (jrc != OXR_EUNSUPPORTED) return jrc;
  }

  const uint64_t n = input->rows;
  FilterShape shape{1, 2};
  if (!use_jit) OXR_TRY(filter_shape(&shape));
  const uint64_t nbatches = n ? (n + batch_rows - 1) / batch_rows : 0;
  // one workgroup per batch walking its tiles in order, with no look-back
  // (knob filter.wg: 0 never, the default; 1 when the batches fill three
  // quarters of the CUs; 2 whenever the query has a generated kernel). On
  // the 200-batch query it measured 1.53 ms at best (8 waves, 8 rows per
  // lane per register set) against 1.31 ms for the tile kernel's look-back:
  // a batch's tiles in series leave each workgroup waiting on its loads.
  const int64_t wg_knob = tune_get("filter.wg", 0);
  const bool use_wg = use_jit && !jf.persistent && jf.wg_fn && !tgt &&
                      (wg_knob >= 2 || (wg_knob == 1 && nbatches * 4 >= (uint64_t)std::max(1, s->dev->cu_count) * 3));
  // the staged kernel (kept rows parked in LDS) unless the last launch of
  // this query shape kept more rows than its stage holds with a margin:
  // tiles past the stage read their later chunks twice (at 50% selectivity
  // 2.2 ms against the tile kernel's 1.7 on the 200M-row table)
  // (inputs under four staged tiles take the tile kernel; knob filter.staged
  // = 2 takes the staged kernel whatever the input)
  bool use_staged = false;
  if (use_jit && jf.staged && !use_wg && !jf.persistent) {
    const double kept = jit_kept_fraction(jf.fn);
    const uint64_t staged_tile = (uint64_t)64 * jf.staged_waves * jf.staged_rows;
    const double share = (double)jf.stage_rows / (double)staged_tile;
    use_staged = tune_get("filter.staged", 1) >= 2 ||
                 (n >= 4 * staged_tile && (kept < 0.0 || kept <= 0.9 * share));
    if (use_staged) {
      // no-allocation calls run what an earlier call compiled, else the tile kernel
      const int lrc = jit_filter_staged(s->dev, &jf, tgt != nullptr);
      if (lrc == OXR_EUNSUPPORTED && tgt) use_staged = false;
      else if (lrc != OXR_OK) return lrc;
    }
  }
  const uint64_t tile_rows = use_wg       ? (uint64_t)64 * jf.wg_waves * jf.wg_rows
                             : use_staged ? (uint64_t)64 * jf.staged_waves * jf.staged_rows
                             : use_jit    ? (uint64_t)64 * jf.waves * jf.rows_per_lane
                                          : (uint64_t)kBlock * shape.v * shape.g;
  std::unique_ptr<Params> guard(new Params());
  Params &P = *guard;
  std::memset(&P, 0, sizeof P);
  P.n = n;
  P.batch_rows = batch_rows;
  P.tiles_per_batch = (batch_rows + tile_rows - 1) / tile_rows;
  P.ntiles = nbatches * P.tiles_per_batch;
  if (P.ntiles >= UINT32_MAX)
    return fail(OXR_EUNSUPPORTED, "rocm.filter", "too many tiles for one launch");
  for (size_t i = 0; i < input->cols.size(); i++) {
    const DevCol &c = input->cols[i];
    P.cols[i].values = c.values;
    P.cols[i].valid = (const uint32_t *)c.validity;
    P.cols[i].type = c.type;
    P.cols[i].dim = c.dim;
    if (n && ((uintptr_t)c.values % 16) != 0)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "column values must be 16-byte aligned");
    if (c.validity && ((uintptr_t)c.validity % 4) != 0)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "validity bitmaps must be 4-byte aligned");
  }
  std::copy(prog.ins.begin(), prog.ins.end(), P.ins);
  std::copy(prog.leaves.begin(), prog.leaves.end(), P.leaves);
  std::copy(prog.bprog.begin(), prog.bprog.end(), P.bprog);
  std::copy(prog.magic.begin(), prog.magic.end(), P.magic);
  for (size_t i = 0; i < prog.slot_cols.size(); i++) P.slot_cols[i] = (int8_t)prog.slot_cols[i];
  P.nslots = (int32_t)prog.slot_cols.size();
  P.nb = (int32_t)prog.bprog.size();
  P.nouts = nproj;

  // scratch: tile status words | error flags | ticket | per-batch counts
  const size_t status_bytes = (size_t)std::max<uint64_t>(P.ntiles, 1) * 8;
  const size_t count_bytes = (size_t)std::max<uint64_t>(nbatches, 1) * 8;
  if (tgt && status_bytes + 64 + count_bytes > tgt->scratch_bytes)
    return fail(OXR_EPLAN, "rocm.filter",
                std::to_string(P.ntiles) + " tiles need " + std::to_string(status_bytes + 64 + count_bytes) +
                    " bytes of scratch; the no-allocation scratch holds " + std::to_string(tgt->scratch_bytes));
  auto release_outputs = [&]() {
    if (!tgt)
      for (auto &c : *out_cols) {
        stream_release(s, c.values);
        stream_release(s, c.validity);
      }
    out_cols->clear();
  };
  for (int i = 0; i < nproj; i++) {
    const ValueProg &v = vps[i];
    DevCol c;
    const int32_t dim = v.col >= 0 ? input->cols[v.col].dim : 0;
    int rc = OXR_OK;
    if (tgt) {
      // the caller's column: the projection's type, room for every input
      // row, and a validity bitmap when the projection can be NULL (its
      // words cleared here: the kernels OR the boundary words in)
      const oxr_column &o = tgt->outs[i];
      if (o.type != v.type || (v.type == OXR_F32_VECTOR && o.dim != dim))
        rc = fail(OXR_EPLAN, "rocm.filter",
                  "output " + std::to_string(i) + " has type " + std::to_string(o.type) + "/" +
                      std::to_string(o.dim) + ", the projection gives " + std::to_string(v.type) + "/" +
                      std::to_string(dim));
      else if (o.len < n)
        rc = fail(OXR_EPLAN, "rocm.filter",
                  "output " + std::to_string(i) + " holds " + std::to_string(o.len) + " rows, the input has " +
                      std::to_string(n));
      else if ((n && !o.values) || (v.nullable && n && !o.validity))
        rc = fail(OXR_EPLAN, "rocm.filter", "output " + std::to_string(i) + " lacks a buffer");
      else if (((uintptr_t)o.values % 16) != 0 || ((uintptr_t)o.validity % 4) != 0)
        rc = fail(OXR_EPLAN, "rocm.filter", "output values must be 16-byte and validity 4-byte aligned");
      c.type = v.type;
      c.dim = dim;
      c.len = n;
      c.values = o.values;
      c.validity = v.nullable ? o.validity : nullptr;
      c.owned = false;
      if (rc == OXR_OK && c.validity) rc = fill_zero(s, c.validity, (n + 31) / 32 * 4);
    } else {
      rc = alloc_column(s, v.type, dim, n, v.nullable, &c);
    }
    if (rc != OXR_OK) {
      release_outputs();
      return rc;
    }
    out_cols->push_back(c);
    out_names->push_back(names && names[i] ? names[i] : "p" + std::to_string(i));
    Out &o = P.outs[i];
    o.values = c.values;
    o.valid = (uint32_t *)c.validity;
    o.vp = v;
    o.dim = dim;
  }

  void *scratch = nullptr;
  int rc = OXR_OK;
  if (tgt) {
    scratch = tgt->scratch;  // its size was checked before any output was touched
  } else {
    rc = stream_alloc(s, status_bytes + 64 + count_bytes, &scratch);
  }
  if (rc != OXR_OK) {
    release_outputs();
    return rc;
  }
  P.status = (uint64_t *)scratch;
  P.result = (uint64_t *)((char *)scratch + status_bytes);
  P.counter = (uint32_t *)((char *)scratch + status_bytes + 16);
  P.counts = (uint64_t *)((char *)scratch + status_bytes + 64);
  rc = fill_zero(s, scratch, status_bytes + 64 + count_bytes);
  if (rc == OXR_OK && P.ntiles && use_jit) {
    JitArgs A;
    std::memset(&A, 0, sizeof A);
    A.n = n;
    A.batch_rows = batch_rows;
    A.tiles_per_batch = P.tiles_per_batch;
    A.status = P.status;
    A.counter = P.counter;
    A.result = P.result;
    A.counts = P.counts;
    for (size_t i = 0; i < input->cols.size(); i++) {
      A.col[i] = input->cols[i].values;
      A.valid[i] = (const uint32_t *)input->cols[i].validity;
    }
    for (int i = 0; i < nproj; i++) {
      A.out[i] = (*out_cols)[i].values;
      A.outv[i] = (uint32_t *)(*out_cols)[i].validity;
      A.vslot[i] = jf.vslot[i];
      A.dim[i] = (*out_cols)[i].dim;
    }
    A.nouts = nproj;
    A.ntiles = P.ntiles;
    std::copy(jf.lits.begin(), jf.lits.end(), A.lit);
    uint64_t grid = P.ntiles;
    if (use_wg) {
      // workgroups per CU (knob filter.wg_grid, default 1); each takes
      // batches blockIdx.x, blockIdx.x + grid, ...
      const uint64_t per_cu = (uint64_t)std::max<int64_t>(1, tune_get("filter.wg_grid", 1));
      grid = std::min<uint64_t>(nbatches, (uint64_t)std::max(1, s->dev->cu_count) * per_cu);
      ProfScope prof(s->s, "filter_project");
      rc = jit_launch(jf.wg_fn, (unsigned)grid, s->s, A, (unsigned)(64 * jf.wg_waves));
    } else if (jf.persistent) {
      // as many workgroups as are resident at once (knob filter.grid: per CU)
      int per_cu = (int)tune_get("filter.grid", 0);
      if (per_cu <= 0) {
        const hipError_t e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, jf.fn, 64 * jf.waves, 0);
        if (e != hipSuccess) rc = fail_hip(e, "occupancy of the generated filter kernel");
        per_cu = std::max(per_cu, 1);
      }
      grid = std::min<uint64_t>(P.ntiles, (uint64_t)std::max(1, s->dev->cu_count) * (uint64_t)per_cu);
    }
    if (rc == OXR_OK && use_staged) {
      ProfScope prof(s->s, "filter_project");
      rc = jit_launch(jf.staged_fn, (unsigned)grid, s->s, A, (unsigned)(64 * jf.staged_waves));
    } else if (rc == OXR_OK && !use_wg) {
      ProfScope prof(s->s, "filter_project");
      rc = jit_launch(jf.fn, (unsigned)grid, s->s, A, (unsigned)(64 * jf.waves));
    }
  } else if (rc == OXR_OK && P.ntiles) {
    ProfScope prof(s->s, "filter_project_vm");
    launch_filter(slot_bucket(P.nslots), shape, dim3((unsigned)P.ntiles), s->s, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "filter_project_kernel launch");
  }
  uint64_t err = 0;
  counts->assign(nbatches, 0);
  std::vector<uint64_t> unpinned;
  if (rc == OXR_OK) {
    // one copy into pinned memory: the error word, then the counts at +64
    // (no-allocation calls: into host memory the call already holds)
    void *host = nullptr;
    if (tgt) {
      unpinned.assign(8 + std::max<uint64_t>(nbatches, 1), 0);
      host = unpinned.data();
    } else {
      rc = stream_pinned(s, 64 + count_bytes, &host);
    }
    if (rc == OXR_OK) {
      hipError_t e = hipMemcpyAsync(host, P.result, 64 + nbatches * 8, hipMemcpyDeviceToHost, s->s);
      if (e == hipSuccess) e = hipStreamSynchronize(s->s);
      if (e != hipSuccess) {
        rc = fail_hip(e, "filter_project result");
      } else {
        std::memcpy(&err, host, 8);
        if (nbatches) std::memcpy(counts->data(), (const char *)host + 64, nbatches * 8);
        if (use_jit && jf.staged && n > 0 && !use_wg && !jf.persistent) {
          uint64_t kept = 0;
          for (uint64_t c : *counts) kept += c;
          jit_note_kept(jf.fn, (double)kept / (double)n);
        }
      }
    }
  }
  if (!tgt) stream_release(s, scratch);
  if (rc == OXR_OK && err) rc = fail(OXR_EEXECUTION, "rocm.filter", "Divide by zero error");
  if (rc != OXR_OK) release_outputs();
  return rc;
}

}  // namespace
}  // namespace oxr

extern "C" int oxr_filter_project(oxr_stream *s, const oxr_batch *input, const oxr_expr *predicate,
                                  const oxr_expr *projection, int32_t nproj,
                                  const char *const *names, oxr_batch **out) {
  clear_error();
  if (!s || !input || !out || (nproj > 0 && !projection))
    return fail(OXR_EPLAN, "rocm.filter", "null argument");
  std::unique_ptr<oxr_batch> ob(new_batch(s, 0));
  std::vector<uint64_t> counts;
  OXR_TRY(run_filter(s, input, std::max<uint64_t>(input->rows, 1), predicate, projection, nproj, names,
                     &ob->cols, &ob->names, &counts));
  ob->rows = counts.empty() ? 0 : counts[0];
  for (auto &c : ob->cols) c.len = ob->rows;
  *out = ob.release();
  return OXR_OK;
}

extern "C" int oxr_filter_project_batches(oxr_stream *s, const oxr_batch *input, uint64_t batch_rows,
                                          const oxr_expr *predicate, const oxr_expr *projection,
                                          int32_t nproj, const char *const *names, oxr_batch **outs,
                                          uint64_t max_batches, uint64_t *nbatches) {
  clear_error();
  if (!s || !input || !outs || !nbatches || (nproj > 0 && !projection))
    return fail(OXR_EPLAN, "rocm.filter", "null argument");
  if (batch_rows == 0 || batch_rows % 32 != 0)
    return fail(OXR_EPLAN, "rocm.filter", "batch_rows must be a non-zero multiple of 32");
  const uint64_t nb = input->rows ? (input->rows + batch_rows - 1) / batch_rows : 0;
  *nbatches = nb;
  if (nb > max_batches)
    return fail(OXR_EPLAN, "rocm.filter",
                "the result has " + std::to_string(nb) + " batches but room for " +
                    std::to_string(max_batches));
  std::vector<DevCol> cols;
  std::vector<std::string> out_names;
  std::vector<uint64_t> counts;
  OXR_TRY(run_filter(s, input, batch_rows, predicate, projection, nproj, names, &cols, &out_names,
                     &counts));
  auto shared = std::make_shared<SharedBuffers>();
  shared->dev = s->dev;
  shared->core = retain(s->core);
  for (const DevCol &c : cols) {
    shared->ptrs.push_back(c.values);
    shared->ptrs.push_back(c.validity);
  }
  for (uint64_t i = 0; i < nb; i++) {
    oxr_batch *b = new oxr_batch();
    b->dev = s->dev;
    b->rows = counts[i];
    b->holders.push_back(shared);
    b->names = out_names;
    for (const DevCol &c : cols) {
      DevCol v = c;
      v.owned = false;
      v.len = counts[i];
      v.values = (char *)c.values + i * batch_rows * type_width(c.type, c.dim);
      v.validity = c.validity ? c.validity + i * batch_rows / 8 : nullptr;
      b->cols.push_back(v);
    }
    outs[i] = b;
  }
  return OXR_OK;
}

// ------------------------------------------------- no-allocation entry point
// The scratch of oxr_filter_project_noalloc: tile look-back words, the tile
// ticket, the error word and the kept-row count, for every call on a device.
// Each device holds its own copy until the process ends. In host code the
// name is only a stand-in for the symbol: kernels get the device address
// hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) uint8_t scratch[OXR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu[kMaxDevices];   // one call at a time per device
void *g_noalloc_ptr[kMaxDevices] = {};  // scratch's address on each device

}  // namespace

extern "C" int oxr_filter_project_noalloc(oxr_stream *s, const oxr_batch *input, const oxr_expr *predicate,
                                          const oxr_expr *projection, int32_t nproj, const oxr_column *outputs,
                                          uint64_t *out_rows) {
  clear_error();
  if (!s || !input || !out_rows || (nproj > 0 && (!projection || !outputs)))
    return fail(OXR_EPLAN, "rocm.filter", "null argument");
  const int dev = s->dev->ordinal;
  if (dev < 0 || dev >= kMaxDevices) return fail(OXR_EDEVICE, "rocm.filter", "device ordinal out of range");
  // held until the call's work has finished: the calls on one device take
  // turns on its scratch
  std::lock_guard<std::mutex> lock(g_noalloc_mu[dev]);
  if (!g_noalloc_ptr[dev]) {
    int prev = 0;
    OXR_HIP(hipGetDevice(&prev));
    if (prev != dev) OXR_HIP(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(no-allocation scratch)");
    g_noalloc_ptr[dev] = ptr;
  }
  FilterTarget tgt;
  tgt.outs = outputs;
  tgt.scratch = g_noalloc_ptr[dev];
  tgt.scratch_bytes = OXR_NOALLOC_SCRATCH_BYTES;
  std::vector<DevCol> cols;
  std::vector<std::string> names;
  std::vector<uint64_t> counts;
  OXR_TRY(run_filter(s, input, std::max<uint64_t>(input->rows, 1), predicate, projection, nproj, nullptr, &cols,
                     &names, &counts, &tgt));
  *out_rows = counts.empty() ? 0 : counts[0];
  return OXR_OK;
}

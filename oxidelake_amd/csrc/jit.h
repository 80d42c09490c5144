// Kernels generated per query shape and compiled at run time with hipRTC
// (host side). The reference ships its CUDA kernels as source and compiles
// them with NVRTC on first use (crates/oxidelake-device/src/cuda/mod.rs:
// 381-431, docs/decisions/ADR-0003-cudarc-dynamic-loading-nvrtc.md); here
// the generated source is specialised to the query's expression structure,
// column types and nullability, so the predicate and projections run as
// straight-line code with its literals as constants (held as kernel
// arguments they would keep ~30 more registers live across the tile's
// look-back: 2.5 ms instead of 1.45 ms for 200M rows). A kernel is compiled
// once per process for each distinct source.
#pragma once

#include <string>
#include <vector>

#include "common.h"
#include "expr.h"
#include "filter_jit.h"
#include "aggregate_jit.h"

namespace oxr {

struct JitFilter {
  hipFunction_t fn = nullptr;
  int rows_per_lane = 0;       // S: a tile is 256 * S rows
  bool persistent = false;     // filter_persistent: a grid that fits at once, tiles by ticket
  int waves = 4;               // W: a workgroup is 64 * W threads
  bool has_wg = false;         // the source holds filter_batch_wg (knob filter.wg set)
  hipFunction_t wg_fn = nullptr;  // filter_batch_wg: one workgroup per batch
  int wg_rows = 8;             // its rows per lane per register set
  int wg_waves = 8;            // its waves per workgroup
  bool staged = false;         // staged_source holds filter_staged (knob filter.staged, default on)
  hipFunction_t staged_fn = nullptr;  // loaded by jit_filter_staged
  int staged_rows = 0;         // its rows per lane (chunk rows x chunks)
  int staged_waves = 8;        // its waves per workgroup
  int stage_rows = 0;          // kept rows its LDS stage holds per tile
  std::vector<int32_t> vslot;  // JitArgs::vslot per projection
  std::vector<uint64_t> lits;  // JitArgs::lit (prepared mode, knob jit.params)
  std::string source;
  std::string staged_source;
};

// Finds or builds the filter kernel of this query shape for `dev`. `vps`
// are the interpreter's compiled projections (their types and nullability
// fix the output columns). OXR_EUNSUPPORTED (no error recorded) means the
// query is outside what generated kernels cover and the interpreter runs it;
// with `cached_only` also that its kernel was not compiled before (nothing
// is compiled or loaded then).
int jit_filter(oxr_device *dev, const std::vector<DevCol> &cols, const oxr_expr *pred,
               const oxr_expr *proj, int32_t nproj, const std::vector<ValueProg> &vps,
               JitFilter *out, bool cached_only = false);

// Launches a generated kernel: `grid` blocks of `block` threads.
int jit_launch(hipFunction_t fn, unsigned grid, hipStream_t s, const JitArgs &args, unsigned block = kJitBlock);
// Loads the staged kernel of a query shape (compiling its module unless
// cached_only, when OXR_EUNSUPPORTED means it is not compiled yet).
int jit_filter_staged(oxr_device *dev, JitFilter *jf, bool cached_only);
// Kept fraction of the last launch of a query shape, by its tile kernel (-1:
// none yet), which picks between the staged and the tile kernel.
double jit_kept_fraction(hipFunction_t fn);
void jit_note_kept(hipFunction_t fn, double fraction);
int jit_launch_raw(hipFunction_t fn, unsigned grid, hipStream_t s, void *args, size_t size,
                   unsigned block = kJitBlock);

// An aggregation shape: the fused filter, the key (a column of the input,
// -1 for one global group, or with `join` the probe key column whose match
// gives the group), and each aggregate's accumulator kind, input expression
// (NULL for COUNT(*)) and whether that input can be NULL.
struct AggJitShape {
  const oxr_expr *filter = nullptr;
  int32_t key_col = -1;
  bool join = false;
  std::vector<uint8_t> kinds;
  std::vector<const oxr_expr *> inputs;
  std::vector<uint8_t> nullable;
  // > 0: generate only the batched-request kernels (aggregate_multi) for
  // batches of up to this many requests; literals are always parameters
  int multi = 0;
};

struct JitAgg {
  hipFunction_t fn = nullptr;
  bool dense = false;  // a plain key: the small-range kernels below exist
  hipFunction_t dense16 = nullptr, dense32 = nullptr;
  bool hot = false;  // the hot-group kernel below exists
  hipFunction_t hot_fn = nullptr;
  bool spill = false;  // rows may spill: the spill's kernels below exist
  hipFunction_t scatter_fn = nullptr, fold_fn = nullptr;
  bool keyless = false;  // a global aggregate: the keyless kernel below exists
  hipFunction_t keyless_fn = nullptr;
  int spill_words = 0;  // 32-bit words per spill record
  int rows_per_lane = 0;
  std::vector<uint64_t> lits;  // AggJitArgs::lit (prepared mode)
  // batched requests (shape.multi): kernels for up to kMultiK[i] requests
  int multi_rows = 0;
  hipFunction_t multi_fn[4] = {}, multi_final[4] = {};
  std::string source;
};

// Request counts the batched-request kernels are instantiated for.
constexpr int kMultiK[4] = {1, 4, 8, 16};
// The largest batch a shape of `naggs` aggregates takes (each request keeps
// its accumulators in registers).
inline int multi_max_requests(int naggs) {
  const int cap = 48 / (naggs > 0 ? naggs : 1);
  int k = 1;
  for (int b : kMultiK)
    if (b <= cap) k = b;
  return k;
}

// Finds or builds the aggregation kernel of this shape for `dev`;
// OXR_EUNSUPPORTED (no error recorded): the interpreter kernel runs it.
int jit_aggregate(oxr_device *dev, const std::vector<DevCol> &cols, const AggJitShape &shape, JitAgg *out);
// Generates the source (and, in prepared mode, the literal values) of this
// shape without compiling or loading anything.
int jit_aggregate_generate(const std::vector<DevCol> &cols, const AggJitShape &shape, JitAgg *out);
// The generated source (compiled for gfx950 when asked), no device needed.
int jit_aggregate_source(const std::vector<DevCol> &cols, const AggJitShape &shape, bool do_compile,
                         std::string *src);

}  // namespace oxr

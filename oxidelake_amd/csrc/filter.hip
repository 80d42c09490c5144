// Fused filter + projection with stream compaction, in one pass.
//
// Replaces `ops::filter_project` (crates/oxidelake-device/src/cuda/ops.rs:
// 143-373), which launches one compare kernel per predicate leaf, one mask
// AND per conjunction, a block count, a device->host copy of the counts with
// a host scan, a scatter of row indices and one gather per projected column.
// Here one kernel per batch:
//   1. each lane evaluates the whole predicate for V=4 consecutive rows in
//      registers (expr.h), loading each column with 16-byte vector loads;
//   2. a wave64 prefix sum of the per-lane counts plus a block scan give each
//      selected row its rank inside the tile;
//   3. a single-pass decoupled look-back over the tiles (tile ids taken from
//      an atomic counter, so every tile a lane waits on is already running)
//      gives the tile's output offset — no host round trip;
//   4. each lane computes the projections of its selected rows and stores
//      them at their compacted positions; output validity bits are assembled
//      in LDS and written as whole words (boundary words with atomicOr).
// Semantics: crates/oxidelake-device/src/cpu/kernels.rs:99-152.
#include <hip/hip_runtime.h>

#include <memory>

#include "common.h"
#include "expr.h"

namespace oxr {
namespace {

constexpr int V = 4;             // consecutive rows per lane
constexpr int kBlock = 256;      // threads per block (4 waves)
constexpr int kGroups = 2;       // row groups per lane
constexpr int kTile = kBlock * V * kGroups;  // rows per tile
constexpr int kWaves = kBlock / 64;
constexpr int kTileWords = kTile / 32 + 2;   // output validity words a tile can touch

constexpr uint64_t kFlagA = 1ull << 62;  // tile aggregate available
constexpr uint64_t kFlagP = 2ull << 62;  // inclusive prefix available
constexpr uint64_t kValMask = (1ull << 62) - 1;

struct Out {
  void *values;
  uint32_t *valid;  // NULL when the projection cannot be NULL
  int32_t type, dim;
  int16_t col;      // >= 0: copy this input column
  uint16_t pstart, plen;
  int16_t pad;
};

struct Params {
  uint64_t n;
  uint64_t ntiles;
  DCol cols[kMaxCols];
  Ins ins[kMaxIns];
  Leaf leaves[kMaxLeaves];
  uint8_t bprog[kMaxBool];
  Out outs[kMaxOuts];
  int32_t nb, nouts;
  uint64_t *status;   // ntiles words, zeroed
  uint32_t *counter;  // tile ticket, zeroed
  uint64_t *result;   // [0] = selected rows, [1] = error flags
};

__device__ inline uint64_t atomic_load_u64(uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_store_u64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ inline uint64_t wave_sum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Decoupled look-back, run by one whole wave. Publishes `agg` for `tile`
// and returns the exclusive prefix of the tiles before it.
__device__ uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t agg, int lane) {
  if (tile == 0) {
    if (lane == 0) atomic_store_u64(&status[0], kFlagP | agg);
    return 0;
  }
  if (lane == 0) atomic_store_u64(&status[tile], kFlagA | agg);
  uint64_t excl = 0;
  int64_t end = (int64_t)tile - 1;  // nearest predecessor in this window
  while (true) {
    const int64_t idx = end - lane;
    uint64_t st = 0;
    if (idx >= 0) {
      // every predecessor took its ticket before this tile did, so it is
      // running and will publish: the wait ends
      do {
        st = atomic_load_u64(&status[idx]);
      } while ((st >> 62) == 0);
    }
    const uint64_t pmask = __ballot(idx >= 0 && (st >> 62) == 2);
    if (pmask) {
      const int first = __builtin_ctzll(pmask);
      uint64_t v = (lane <= first && idx >= 0) ? (st & kValMask) : 0;
      excl += wave_sum64(v);
      break;
    }
    excl += wave_sum64(idx >= 0 ? (st & kValMask) : 0);
    end -= 64;
  }
  if (lane == 0) atomic_store_u64(&status[tile], kFlagP | (excl + agg));
  return excl;
}

__device__ inline void store_value(void *base, int type, uint64_t pos, uint64_t v) {
  if (type == OXR_INT32 || type == OXR_FLOAT32) ((uint32_t *)base)[pos] = (uint32_t)v;
  else ((uint64_t *)base)[pos] = v;
}

__global__ __launch_bounds__(kBlock) void filter_project_kernel(const Params p) {
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
  const uint64_t tile = s_tile;
  const uint64_t tile_base = tile * (uint64_t)kTile;
  bool div0 = false;

  // 1. predicate
  uint32_t sel[kGroups];
  uint32_t excl[kGroups];
#pragma unroll
  for (int j = 0; j < kGroups; j++) {
    const uint64_t row0 = tile_base + ((uint64_t)(j * kBlock + tid)) * V;
    sel[j] = eval_predicate<V>(p.leaves, p.bprog, p.nb, p.ins, p.cols, row0, p.n, &div0);
    const uint32_t cnt = __builtin_popcount(sel[j]);
    const uint32_t incl = wave_incl_scan(cnt, lane);
    excl[j] = incl - cnt;
    if (lane == 63) s_wtot[j][wave] = incl;
  }
  __syncthreads();
  // 2. block scan over (group, wave) in row order
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
  // 3. look-back
  if (wave == 0) {
    const uint64_t pre = lookback(p.status, tile, s_total, lane);
    if (lane == 0) {
      s_prefix = pre;
      if (tile == p.ntiles - 1) p.result[0] = pre + s_total;
    }
  }
  __syncthreads();
  const uint64_t prefix = s_prefix;
  const uint64_t word0 = prefix >> 5;

  // 4. projections
  for (int o = 0; o < p.nouts; o++) {
    const Out out = p.outs[o];
#pragma unroll
    for (int j = 0; j < kGroups; j++) {
      const uint64_t row0 = tile_base + ((uint64_t)(j * kBlock + tid)) * V;
      uint64_t pos = prefix + s_woff[j][wave] + excl[j];
      if (out.type == OXR_F32_VECTOR) {
        // whole-wave copy of each selected row's dim floats
        const int dim = out.dim;
        const float *src = (const float *)p.cols[out.col].values;
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
        if (out.valid) {
          uint64_t x[V];
          uint32_t vm;
          load_rows<V>(p.cols[out.col], row0, p.n, x, vm);
          (void)x;
          uint64_t at = pos;
          for (int q = 0; q < V; q++) {
            if (!((sel[j] >> q) & 1)) continue;
            if ((vm >> q) & 1) atomicOr(&s_bits[o][(at >> 5) - word0], 1u << (at & 31));
            at++;
          }
        }
        continue;
      }
      if (!sel[j] && !out.valid) continue;
      uint64_t v[V];
      uint32_t vm = 0;
      if (out.col >= 0) {
        load_rows<V>(p.cols[out.col], row0, p.n, v, vm);
      } else {
        uint64_t r[kRegs][V];
        uint32_t m[kRegs] = {0, 0, 0, 0};
        run_program<V>(p.ins, out.pstart, out.plen, p.cols, row0, p.n, r, m, &div0);
#pragma unroll
        for (int q = 0; q < V; q++) v[q] = r[0][q];
        vm = m[0];
      }
#pragma unroll
      for (int q = 0; q < V; q++) {
        if (!((sel[j] >> q) & 1)) continue;
        store_value(out.values, out.type, pos, v[q]);
        if (out.valid && ((vm >> q) & 1))
          atomicOr(&s_bits[o][(pos >> 5) - word0], 1u << (pos & 31));
        pos++;
      }
    }
  }
  if (div0) atomicOr((unsigned long long *)&p.result[1], 1ull);
  // 5. validity words
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

}  // namespace
}  // namespace oxr

using namespace oxr;

extern "C" int oxr_filter_project(oxr_stream *s, const oxr_batch *input, const oxr_expr *predicate,
                                  const oxr_expr *projection, int32_t nproj,
                                  const char *const *names, oxr_batch **out) {
  clear_error();
  if (!s || !input || !out || (nproj > 0 && !projection))
    return fail(OXR_EPLAN, "rocm.filter", "null argument");
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

  const uint64_t n = input->rows;
  Params *pp = new Params();
  std::unique_ptr<Params> guard(pp);
  Params &P = *pp;
  std::memset(&P, 0, sizeof P);
  P.n = n;
  P.ntiles = (n + kTile - 1) / kTile;
  for (size_t i = 0; i < input->cols.size(); i++) {
    const DevCol &c = input->cols[i];
    P.cols[i].values = c.values;
    P.cols[i].valid = (const uint32_t *)c.validity;
    P.cols[i].type = c.type;
    P.cols[i].dim = c.dim;
    const uintptr_t a = (uintptr_t)c.values;
    if (n && (a % 16) != 0)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "column values must be 16-byte aligned");
    if (c.validity && ((uintptr_t)c.validity % 4) != 0)
      return fail(OXR_EUNSUPPORTED, "rocm.filter", "validity bitmaps must be 4-byte aligned");
  }
  std::copy(prog.ins.begin(), prog.ins.end(), P.ins);
  std::copy(prog.leaves.begin(), prog.leaves.end(), P.leaves);
  std::copy(prog.bprog.begin(), prog.bprog.end(), P.bprog);
  P.nb = (int32_t)prog.bprog.size();
  P.nouts = nproj;

  std::unique_ptr<oxr_batch> ob(new_batch(s, 0));
  auto release_outputs = [&]() {
    for (auto &c : ob->cols) {
      stream_release(s, c.values);
      stream_release(s, c.validity);
    }
    ob->cols.clear();
  };
  for (int i = 0; i < nproj; i++) {
    const ValueProg &v = vps[i];
    DevCol c;
    const int32_t dim = v.col >= 0 ? input->cols[v.col].dim : 0;
    int rc = alloc_column(s, v.type, dim, n, v.nullable, &c);
    if (rc != OXR_OK) {
      release_outputs();
      return rc;
    }
    ob->cols.push_back(c);
    ob->names.push_back(names && names[i] ? names[i] : "p" + std::to_string(i));
    Out &o = P.outs[i];
    o.values = c.values;
    o.valid = (uint32_t *)c.validity;
    o.type = v.type;
    o.dim = dim;
    o.col = v.col;
    o.pstart = v.start;
    o.plen = v.len;
  }

  // scratch: tile status words, ticket, result
  const size_t status_bytes = (size_t)std::max<uint64_t>(P.ntiles, 1) * 8;
  void *scratch = nullptr;
  int rc = stream_alloc(s, status_bytes + 64, &scratch);
  if (rc != OXR_OK) {
    release_outputs();
    return rc;
  }
  P.status = (uint64_t *)scratch;
  P.result = (uint64_t *)((char *)scratch + status_bytes);
  P.counter = (uint32_t *)((char *)scratch + status_bytes + 16);
  rc = fill_zero(s, scratch, status_bytes + 64);
  uint64_t host_result[2] = {0, 0};
  if (rc == OXR_OK && P.ntiles) {
    ProfScope prof(s->s, "filter_project");
    hipLaunchKernelGGL(filter_project_kernel, dim3((unsigned)P.ntiles), dim3(kBlock), 0, s->s, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail_hip(e, "filter_project_kernel launch");
  }
  if (rc == OXR_OK) {
    hipError_t e = hipMemcpyAsync(host_result, P.result, 16, hipMemcpyDeviceToHost, s->s);
    if (e == hipSuccess) e = hipStreamSynchronize(s->s);
    if (e != hipSuccess) rc = fail_hip(e, "filter_project result");
  }
  stream_release(s, scratch);
  if (rc != OXR_OK) {
    release_outputs();
    return rc;
  }
  if (host_result[1]) {
    release_outputs();
    return fail(OXR_EEXECUTION, "rocm.filter", "Divide by zero error");
  }
  ob->rows = host_result[0];
  for (auto &c : ob->cols) c.len = ob->rows;
  *out = ob.release();
  return OXR_OK;
}

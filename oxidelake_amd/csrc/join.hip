// Inner hash join on one Int64 (or Int32) key: build, probe, gather.
//
// Replaces `ops::prepare_join_build` / `probe_join` / `hash_join`
// (crates/oxidelake-device/src/cuda/ops.rs:379-621) and kernels/cuda/
// hash_join.cu. The reference builds its index on the host (JoinSlots::build,
// join_slots.rs:64-122) and scans each probe batch's match counts on the
// host. Here the table is built on the device:
//   1. a min/max/count reduction of the build keys picks the dense
//      (direct-address) or the hashed form (join.h);
//   2. one pass claims every key's slot and counts its rows;
//   3. an exclusive scan of the counts gives each run's start;
//   4. rows are placed in build order within a key: directly when every key
//      is unique, else by a stable radix sort of (slot, row) pairs.
// A probe looks every row up once and scans the match counts on the device
// (u64 offsets, as JoinMatches keeps them, kernels.rs:294-342); a gather
// turns any range of matches into output rows (probe-row order, then build
// order: backend.rs:148-152), so output can be cut into bounded chunks.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <memory>

#include "common.h"
#include "join.h"

namespace oxr {
namespace {

struct KeyCol {
  const void *values;
  const uint32_t *valid;
  int32_t type;
};

__device__ inline bool key_at(const KeyCol &k, uint64_t i, int64_t *out) {
  if (k.valid && !((k.valid[i >> 5] >> (i & 31)) & 1)) return false;
  *out = k.type == OXR_INT32 ? (int64_t)((const int32_t *)k.values)[i] : ((const int64_t *)k.values)[i];
  return true;
}

// [0] = min, [1] = max, [2] = non-NULL keys
__global__ void key_stats_kernel(KeyCol k, uint64_t n, int64_t *out) {
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  unsigned long long cnt = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int64_t v;
    if (key_at(k, i, &v)) {
      mn = v < mn ? v : mn;
      mx = v > mx ? v : mx;
      cnt++;
    }
  }
  for (int d = 32; d > 0; d >>= 1) {
    int64_t a = __shfl_xor(mn, d, 64), b = __shfl_xor(mx, d, 64);
    unsigned long long c = __shfl_xor(cnt, d, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
    cnt += c;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin((long long *)&out[0], (long long)mn);
    atomicMax((long long *)&out[1], (long long)mx);
    atomicAdd((unsigned long long *)&out[2], cnt);
  }
}

__global__ void init_stats_kernel(int64_t *out) {
  out[0] = INT64_MAX;
  out[1] = INT64_MIN;
  out[2] = 0;
  out[3] = 0;
}

__global__ void fill_keys_kernel(int64_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = kJoinEmpty;
}

// Claims each key's slot and counts its rows; row_slot[j] = slot or UINT32_MAX.
__global__ void claim_kernel(JTable t, KeyCol k, uint64_t n, uint32_t *row_slot,
                             unsigned long long *distinct, uint32_t *overflow) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  int64_t key;
  if (!key_at(k, j, &key)) {
    row_slot[j] = UINT32_MAX;
    return;
  }
  uint64_t slot;
  if (t.dense) {
    slot = (uint64_t)key - (uint64_t)t.min_key;
  } else if (key == kJoinEmpty) {
    slot = t.slots;
  } else {
    const uint64_t mask = t.slots - 1;
    uint64_t h = join_hash((uint64_t)key) & mask;
    slot = UINT64_MAX;
    for (uint64_t p = 0; p < t.slots; p++) {
      int64_t cur = __hip_atomic_load(&t.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == key) {
        slot = h;
        break;
      }
      if (cur == kJoinEmpty) {
        unsigned long long old = atomicCAS((unsigned long long *)&t.keys[h],
                                           (unsigned long long)kJoinEmpty, (unsigned long long)key);
        if ((int64_t)old == kJoinEmpty || (int64_t)old == key) {
          slot = h;
          break;
        }
      }
      h = (h + 1) & mask;
    }
    if (slot == UINT64_MAX) {
      atomicOr(overflow, 1u);
      row_slot[j] = UINT32_MAX;
      return;
    }
  }
  const uint32_t before = atomicAdd(&t.count[slot], 1u);
  if (before == 0) atomicAdd(distinct, 1ull);
  row_slot[j] = (uint32_t)slot;
}

// Unique keys: each row is its run.
__global__ void place_unique_kernel(JTable t, const uint32_t *row_slot, uint64_t n) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t s = row_slot[j];
  if (s != UINT32_MAX) t.rows[t.start[s]] = (uint32_t)j;
}

__global__ void iota_kernel(uint32_t *p, uint64_t n) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j < n) p[j] = (uint32_t)j;
}

__global__ void probe_kernel(JTable t, KeyCol k, uint64_t n, uint32_t *starts, uint64_t *counts) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    counts[n] = 0;
    return;
  }
  int64_t key;
  uint64_t slot = UINT64_MAX;
  if (key_at(k, i, &key)) slot = join_find(t, key);
  if (slot == UINT64_MAX) {
    starts[i] = 0;
    counts[i] = 0;
  } else {
    starts[i] = t.start[slot];
    counts[i] = t.count[slot];
  }
}

constexpr int kMaxGather = 32;

struct GCol {
  const void *src;
  const uint32_t *src_valid;
  void *dst;
  uint32_t *dst_valid;
  int32_t width;   // bytes per row
  int32_t build;   // 0 = probe side, 1 = build side
};

struct GatherParams {
  JTable t;
  const uint64_t *offsets;
  const uint32_t *starts;
  uint64_t probe_rows;
  uint64_t first, rows;
  int32_t ncols;
  GCol cols[kMaxGather];
};

__global__ void gather_kernel(const GatherParams p) {
  const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= p.rows) return;
  const uint64_t m = p.first + j;
  // probe row holding match m: the last whose offset is <= m
  uint64_t lo = 0, hi = p.probe_rows + 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (p.offsets[mid] <= m) lo = mid + 1;
    else hi = mid;
  }
  const uint64_t i = lo - 1;
  const uint64_t b = p.t.rows[p.starts[i] + (m - p.offsets[i])];
  for (int c = 0; c < p.ncols; c++) {
    const GCol g = p.cols[c];
    const uint64_t r = g.build ? b : i;
    if (g.width == 8) ((uint64_t *)g.dst)[j] = ((const uint64_t *)g.src)[r];
    else if (g.width == 4) ((uint32_t *)g.dst)[j] = ((const uint32_t *)g.src)[r];
    else {
      const uint32_t *s = (const uint32_t *)((const char *)g.src + r * (uint64_t)g.width);
      uint32_t *d = (uint32_t *)((char *)g.dst + j * (uint64_t)g.width);
      for (int w = 0; w < g.width / 4; w++) d[w] = s[w];
    }
    if (g.dst_valid) {
      const bool v = !g.src_valid || ((g.src_valid[r >> 5] >> (r & 31)) & 1);
      if (v) atomicOr(&g.dst_valid[j >> 5], 1u << (j & 31));
    }
  }
}

unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + 255) / 256); }

int key_col(const oxr_batch *b, int32_t key, KeyCol *out, const char *what) {
  if (key < 0 || key >= (int32_t)b->cols.size())
    return fail(OXR_EPLAN, "rocm.join", std::string(what) + " key column out of range");
  const DevCol &c = b->cols[key];
  if (c.type != OXR_INT64 && c.type != OXR_INT32)
    return fail(OXR_EUNSUPPORTED, "cpu.int64_column",
                std::string(what) + " key column has a type other than Int64/Int32");
  out->values = c.values;
  out->valid = (const uint32_t *)c.validity;
  out->type = c.type;
  return OXR_OK;
}

}  // namespace
}  // namespace oxr

using namespace oxr;

extern "C" {

int oxr_join_build(oxr_stream *s, const oxr_batch *build, int32_t key, oxr_join_table **out) {
  clear_error();
  if (!s || !build || !out) return fail(OXR_EPLAN, "rocm.join", "null argument");
  KeyCol k;
  OXR_TRY(key_col(build, key, &k, "build"));
  const uint64_t n = build->rows;
  if (n >= UINT32_MAX) return fail(OXR_EEXECUTION, "rocm.join", "build side exceeds u32::MAX rows");
  // the device scan and sort of the build side take 32-bit signed counts
  // (slots are at most 2 * n rounded up to a power of two, so n < 2^29 keeps
  // them below 2^31)
  if (n >= (1ull << 29))
    return fail(OXR_EUNSUPPORTED, "rocm.join", "build side of 2^29 rows or more");
  std::unique_ptr<oxr_join_table> jt(new oxr_join_table());
  static std::atomic<uint64_t> next_serial{0};
  jt->serial = ++next_serial;
  jt->dev = s->dev;
  jt->owner = retain(s->core);
  jt->build_rows = n;
  jt->key_type = k.type;
  // 1. key statistics
  void *scratch = nullptr;
  OXR_TRY(stream_alloc(s, 64, &scratch));
  int64_t *stats = (int64_t *)scratch;
  hipLaunchKernelGGL(init_stats_kernel, dim3(1), dim3(1), 0, s->s, stats);
  if (n)
    hipLaunchKernelGGL(key_stats_kernel, dim3(std::min<unsigned>(grid_for(n), 2048)), dim3(256), 0, s->s,
                       k, n, stats);
  int64_t hs[4];
  hipError_t e = hipMemcpyAsync(hs, stats, 32, hipMemcpyDeviceToHost, s->s);
  if (e == hipSuccess) e = hipStreamSynchronize(s->s);
  stream_release(s, scratch);
  if (e != hipSuccess) return fail_hip(e, "join build statistics");
  const uint64_t nonnull = (uint64_t)hs[2];
  const int64_t mn = hs[0], mx = hs[1];
  // 2. table form
  JTable &t = jt->t;
  std::memset(&t, 0, sizeof t);
  const uint64_t range = nonnull ? (uint64_t)mx - (uint64_t)mn + 1 : 0;
  const bool dense = nonnull && range != 0 && range <= std::max<uint64_t>(2 * nonnull, 64) &&
                     range < (1ull << 31);
  t.dense = dense;
  if (dense) {
    t.min_key = mn;
    t.slots = range;
  } else {
    uint64_t cap = 16;
    while (cap < 2 * nonnull) cap <<= 1;
    if (cap > (1ull << 31))
      return fail(OXR_EUNSUPPORTED, "device.join_table",
                  std::to_string(nonnull) + " build keys exceed the device index");
    t.slots = cap;
  }
  const uint64_t slots = t.slots + 1;
  size_t bytes = slots * 8 /*start+count*/ + (dense ? 0 : slots * 8) + n * 4 + 64;
  OXR_TRY(stream_alloc(s, bytes, &jt->mem));
  char *p = (char *)jt->mem;
  t.start = (uint32_t *)p;
  p += slots * 4;
  t.count = (uint32_t *)p;
  p += slots * 4;
  p = (char *)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  if (!dense) {
    t.keys = (int64_t *)p;
    p += slots * 8;
  }
  uint32_t *row_slot = (uint32_t *)p;
  OXR_TRY(stream_alloc(s, std::max<uint64_t>(nonnull, 1) * 4, &jt->rows_mem));
  t.rows = (uint32_t *)jt->rows_mem;
  void *cnt_mem = nullptr;
  OXR_TRY(stream_alloc(s, 64, &cnt_mem));
  unsigned long long *distinct = (unsigned long long *)cnt_mem;
  uint32_t *overflow = (uint32_t *)((char *)cnt_mem + 8);
  int rc = fill_zero(s, cnt_mem, 64);
  if (rc == OXR_OK) rc = fill_zero(s, t.count, slots * 4);
  if (rc == OXR_OK && !dense)
    hipLaunchKernelGGL(fill_keys_kernel, dim3(std::min<unsigned>(grid_for(slots), 8192)), dim3(256), 0,
                       s->s, t.keys, slots);
  // 3. claim + count
  if (rc == OXR_OK && n) {
    ProfScope prof(s->s, "join_build");
    hipLaunchKernelGGL(claim_kernel, dim3(grid_for(n)), dim3(256), 0, s->s, t, k, n, row_slot,
                       distinct, overflow);
  }
  uint64_t host[2] = {0, 0};
  if (rc == OXR_OK) {
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, cnt_mem, 16, hipMemcpyDeviceToHost, s->s);
    if (e != hipSuccess) rc = fail_hip(e, "join claim");
  }
  // 4. starts: exclusive scan of the counts
  void *tmp = nullptr;
  if (rc == OXR_OK) {
    size_t tb = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, t.count, t.start, (int)slots, s->s);
    if (e == hipSuccess) rc = stream_alloc(s, tb + 16, &tmp);
    if (rc == OXR_OK && e == hipSuccess)
      e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, t.count, t.start, (int)slots, s->s);
    if (rc == OXR_OK && e == hipSuccess) e = hipStreamSynchronize(s->s);
    if (rc == OXR_OK && e != hipSuccess) rc = fail_hip(e, "join start scan");
  }
  stream_release(s, tmp);
  tmp = nullptr;
  if (rc == OXR_OK && (uint32_t)host[1])
    rc = fail(OXR_EEXECUTION, "rocm.join", "device join index is full");
  const uint64_t dist = host[0];
  const bool unique = dist == nonnull;
  t.unique = unique;
  // 5. rows in build order within a key
  if (rc == OXR_OK && nonnull) {
    if (unique) {
      hipLaunchKernelGGL(place_unique_kernel, dim3(grid_for(n)), dim3(256), 0, s->s, t, row_slot, n);
      e = hipGetLastError();
      if (e != hipSuccess) rc = fail_hip(e, "place_unique_kernel");
    } else {
      void *sort_mem = nullptr;
      rc = stream_alloc(s, n * 4 * 3 + 64, &sort_mem);
      if (rc == OXR_OK) {
        uint32_t *iota = (uint32_t *)sort_mem;
        uint32_t *keys_out = iota + n;
        uint32_t *rows_out = keys_out + n;
        hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(256), 0, s->s, iota, n);
        // a stable sort on all 32 bits: NULL rows (UINT32_MAX) sort last
        size_t tb = 0;
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, row_slot, keys_out, iota, rows_out, (int)n,
                                               0, 32, s->s);
        if (e == hipSuccess) rc = stream_alloc(s, tb + 16, &tmp);
        if (rc == OXR_OK && e == hipSuccess)
          e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, row_slot, keys_out, iota, rows_out, (int)n, 0,
                                                 32, s->s);
        if (rc == OXR_OK && e == hipSuccess)
          e = hipMemcpyAsync(t.rows, rows_out, nonnull * 4, hipMemcpyDeviceToDevice, s->s);
        if (rc == OXR_OK && e == hipSuccess) e = hipStreamSynchronize(s->s);
        if (rc == OXR_OK && e != hipSuccess) rc = fail_hip(e, "join row placement");
        stream_release(s, tmp);
        stream_release(s, sort_mem);
      }
    }
  }
  if (rc == OXR_OK) {
    e = hipStreamSynchronize(s->s);
    if (e != hipSuccess) rc = fail_hip(e, "join build");
  }
  stream_release(s, cnt_mem);
  if (rc != OXR_OK) {
    stream_release(s, jt->mem);
    stream_release(s, jt->rows_mem);
    return rc;
  }
  jt->distinct = dist;
  jt->nrows = nonnull;
  *out = jt.release();
  return OXR_OK;
}

int oxr_join_table_info(const oxr_join_table *t, uint64_t *distinct_keys, uint64_t *rows, int *dense) {
  if (!t) return fail(OXR_EPLAN, "rocm.join", "null table");
  if (distinct_keys) *distinct_keys = t->distinct;
  if (rows) *rows = t->nrows;
  if (dense) *dense = t->t.dense;
  return OXR_OK;
}

void oxr_join_table_free(oxr_join_table *t) {
  if (!t) return;
  t->dev->arena->release(t->mem, t->owner->s);
  t->dev->arena->release(t->rows_mem, t->owner->s);
  release(t->owner);
  delete t;
}

int oxr_join_probe(oxr_stream *s, const oxr_join_table *table, const oxr_batch *probe, int32_t key,
                   oxr_join_matches **out, uint64_t *total) {
  clear_error();
  if (!s || !table || !probe || !out) return fail(OXR_EPLAN, "rocm.join", "null argument");
  KeyCol k;
  OXR_TRY(key_col(probe, key, &k, "probe"));
  const uint64_t n = probe->rows;
  if (n >= UINT32_MAX) return fail(OXR_EEXECUTION, "rocm.join", "batch exceeds u32::MAX rows");
  // the device scan of the n + 1 match counts takes a 32-bit signed count
  if (n + 1 > (uint64_t)INT32_MAX)
    return fail(OXR_EUNSUPPORTED, "rocm.join", "probe batch of 2^31 - 1 rows or more");
  std::unique_ptr<oxr_join_matches> m(new oxr_join_matches());
  m->dev = s->dev;
  m->owner = retain(s->core);
  m->probe_rows = n;
  m->table = table;
  // offsets (n+1) | counts (n+1) | starts (n)
  const size_t ob = (n + 1) * 8;
  OXR_TRY(stream_alloc(s, ob * 2 + n * 4 + 64, &m->mem));
  m->offsets = (uint64_t *)m->mem;
  uint64_t *counts = (uint64_t *)((char *)m->mem + ob);
  m->starts = (uint32_t *)((char *)m->mem + 2 * ob);
  int rc = OXR_OK;
  {
    ProfScope prof(s->s, "join_probe");
    hipLaunchKernelGGL(probe_kernel, dim3(grid_for(n + 1)), dim3(256), 0, s->s, table->t, k, n,
                       m->starts, counts);
  }
  hipError_t e = hipGetLastError();
  void *tmp = nullptr;
  if (e == hipSuccess) {
    size_t tb = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts, m->offsets, (int)(n + 1), s->s);
    if (e == hipSuccess) rc = stream_alloc(s, tb + 16, &tmp);
    if (rc == OXR_OK && e == hipSuccess)
      e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, m->offsets, (int)(n + 1), s->s);
  }
  uint64_t tot = 0;
  uint64_t *pin = nullptr;  // the match total, read back through pinned memory
  if (rc == OXR_OK && e == hipSuccess) rc = stream_pinned(s, 8, (void **)&pin);
  if (rc == OXR_OK && e == hipSuccess)
    e = hipMemcpyAsync(pin, m->offsets + n, 8, hipMemcpyDeviceToHost, s->s);
  if (rc == OXR_OK && e == hipSuccess) e = hipStreamSynchronize(s->s);
  if (rc == OXR_OK && e == hipSuccess) tot = *pin;
  stream_release(s, tmp);
  if (rc == OXR_OK && e != hipSuccess) rc = fail_hip(e, "join probe");
  if (rc != OXR_OK) {
    stream_release(s, m->mem);
    return rc;
  }
  m->total = tot;
  if (total) *total = tot;
  *out = m.release();
  return OXR_OK;
}

void oxr_join_matches_free(oxr_join_matches *m) {
  if (!m) return;
  m->dev->arena->release(m->mem, m->owner->s);
  release(m->owner);
  delete m;
}

int oxr_join_gather(oxr_stream *s, const oxr_batch *probe, const oxr_batch *build,
                    const oxr_join_table *table, const oxr_join_matches *matches, uint64_t first,
                    uint64_t rows, const int32_t *projection, int32_t nproj, oxr_batch **out) {
  clear_error();
  if (!s || !probe || !build || !table || !matches || !out)
    return fail(OXR_EPLAN, "rocm.join", "null argument");
  if (matches->probe_rows != probe->rows)
    return fail(OXR_EEXECUTION, "rocm.join",
                "join matches cover " + std::to_string(matches->probe_rows) +
                    " probe rows but the probe batch has " + std::to_string(probe->rows));
  if (matches->table != table)
    return fail(OXR_EEXECUTION, "rocm.join", "join matches were probed against another table");
  if (table->build_rows != build->rows)
    return fail(OXR_EEXECUTION, "rocm.join", "the build batch is not the one the table indexes");
  if (first + rows < first || first + rows > matches->total)
    return fail(OXR_EEXECUTION, "rocm.join",
                "matches " + std::to_string(first) + "..+" + std::to_string(rows) +
                    " are outside a probe with " + std::to_string(matches->total) + " matches");
  const int32_t width = (int32_t)probe->cols.size();
  const int32_t all = width + (int32_t)build->cols.size();
  std::vector<int32_t> sel;
  if (nproj < 0) {
    for (int32_t i = 0; i < all; i++) sel.push_back(i);
  } else {
    sel.assign(projection, projection + nproj);
  }
  if ((int)sel.size() > kMaxGather)
    return fail(OXR_EUNSUPPORTED, "rocm.join", "too many output columns");
  std::unique_ptr<GatherParams> gp(new GatherParams());
  GatherParams &P = *gp;
  std::memset(&P, 0, sizeof P);
  P.t = table->t;
  P.offsets = matches->offsets;
  P.starts = matches->starts;
  P.probe_rows = matches->probe_rows;
  P.first = first;
  P.rows = rows;
  P.ncols = (int32_t)sel.size();
  std::unique_ptr<oxr_batch> ob(new_batch(s, rows));
  auto cleanup = [&]() {
    for (auto &c : ob->cols) {
      stream_release(s, c.values);
      stream_release(s, c.validity);
    }
  };
  for (size_t c = 0; c < sel.size(); c++) {
    const int32_t idx = sel[c];
    if (idx < 0 || idx >= all) {
      cleanup();
      return fail(OXR_EPLAN, "rocm.join",
                  "join projection index " + std::to_string(idx) + " out of range for " +
                      std::to_string(all) + " columns");
    }
    const bool b = idx >= width;
    const DevCol &src = b ? build->cols[idx - width] : probe->cols[idx];
    const std::string &name = b ? build->names[idx - width] : probe->names[idx];
    DevCol d;
    int rc = alloc_column(s, src.type, src.dim, rows, src.validity != nullptr, &d);
    if (rc != OXR_OK) {
      cleanup();
      return rc;
    }
    ob->cols.push_back(d);
    ob->names.push_back(name);
    GCol &g = P.cols[c];
    g.src = src.values;
    g.src_valid = (const uint32_t *)src.validity;
    g.dst = d.values;
    g.dst_valid = (uint32_t *)d.validity;
    g.width = (int32_t)type_width(src.type, src.dim);
    g.build = b;
  }
  if (rows) {
    ProfScope prof(s->s, "join_gather");
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(rows)), dim3(256), 0, s->s, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      cleanup();
      return fail_hip(e, "join gather launch");
    }
  }
  hipError_t e = hipStreamSynchronize(s->s);
  if (e != hipSuccess) {
    cleanup();
    return fail_hip(e, "join gather");
  }
  *out = ob.release();
  return OXR_OK;
}

}  // extern "C"

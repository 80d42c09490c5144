// Device, HBM arena, streams, buffers and batches, plus the Arrow C Data
// Interface import (upload) and export (download).
//
// Replaces the CUDA backend's plumbing: CudaBackend::probe/new
// (crates/oxidelake-device/src/cuda/mod.rs:293-365) and upload/download
// (cuda/mod.rs:445-498, 620-661), with the transfer format of
// crates/oxidelake-device/src/columns.rs:25-309.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <memory>

#include "common.h"

namespace oxr {

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string t_error;
thread_local std::string t_feature;
}  // namespace

int fail(int code, const char *feature, const std::string &msg) {
  t_error = msg;
  t_feature = feature ? feature : "";
  return code;
}

int fail_hip(hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(OXR_EALLOCATION, "rocm.alloc", msg);
  return fail(OXR_EDEVICE, "rocm.device", msg);
}

void clear_error() {
  t_error.clear();
  t_feature.clear();
}

// ------------------------------------------------------------------- arena
size_t Arena::round_size(size_t bytes) {
  if (bytes < 256) return 256;
  if (bytes <= (1u << 20)) {  // powers of two up to 1 MiB
    size_t s = 256;
    while (s < bytes) s <<= 1;
    return s;
  }
  const size_t g = 2u << 20;  // 2 MiB granules above
  return (bytes + g - 1) / g * g;
}

Arena::~Arena() { trim(); }

int Arena::alloc(size_t bytes, hipStream_t s, void **out) {
  const size_t sz = round_size(bytes);
  Block found{nullptr, nullptr, nullptr};
  size_t found_size = 0;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(sz);
    // best fit, at most 1/8 larger than asked (small classes match exactly)
    if (it != free_.end() && it->first <= sz + sz / 8) {
      found = it->second;
      found_size = it->first;
      free_.erase(it);
      cached_ -= found_size;
      live_[found.ptr] = found_size;
      in_use_ += found_size;
    }
  }
  if (found.ptr) {
    if (found.stream != s) {
      // order this stream's use after the work that last touched the block
      hipError_t e = hipStreamWaitEvent(s, found.freed, 0);
      if (e != hipSuccess) {
        release(found.ptr, s);
        return fail_hip(e, "hipStreamWaitEvent");
      }
    }
    hipEventDestroy(found.freed);
    *out = found.ptr;
    return OXR_OK;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    // give cached blocks back to the driver and retry once
    trim();
    e = hipMalloc(&p, sz);
    if (e != hipSuccess)
      return fail(OXR_EALLOCATION, "rocm.alloc",
                  "hipMalloc of " + std::to_string(sz) + " bytes failed: " + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> g(mu_);
  live_[p] = sz;
  in_use_ += sz;
  *out = p;
  return OXR_OK;
}

void Arena::release(void *p, hipStream_t s) {
  if (!p) return;
  hipEvent_t ev = nullptr;
  hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  hipEventRecord(ev, s);
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) {
    hipEventDestroy(ev);
    return;
  }
  size_t sz = it->second;
  live_.erase(it);
  in_use_ -= sz;
  free_.emplace(sz, Block{p, s, ev});
  cached_ += sz;
}

void Arena::stats(uint64_t *in_use, uint64_t *cached) {
  std::lock_guard<std::mutex> g(mu_);
  if (in_use) *in_use = in_use_;
  if (cached) *cached = cached_;
}

void Arena::trim() {
  std::multimap<size_t, Block> blocks;
  {
    std::lock_guard<std::mutex> g(mu_);
    blocks.swap(free_);
    cached_ = 0;
  }
  for (auto &kv : blocks) {
    hipEventSynchronize(kv.second.freed);
    hipEventDestroy(kv.second.freed);
    hipFree(kv.second.ptr);
  }
}

// ------------------------------------------------------------------ helpers
size_t type_width(int32_t type, int32_t dim) {
  switch (type) {
    case OXR_INT32:
    case OXR_FLOAT32: return 4;
    case OXR_INT64:
    case OXR_FLOAT64: return 8;
    case OXR_F32_VECTOR: return (size_t)4 * (size_t)(dim > 0 ? dim : 0);
  }
  return 0;
}

oxr_device *device_of(oxr_stream *s) { return s->dev; }

int stream_alloc(oxr_stream *s, size_t bytes, void **out) {
  return s->dev->arena->alloc(bytes, s->s, out);
}

void stream_release(oxr_stream *s, void *p) { s->dev->arena->release(p, s->s); }

__global__ void zero_fill_kernel(uint4 *p, size_t n16, uint8_t *tail, size_t ntail) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

int fill_zero(oxr_stream *s, void *p, size_t bytes) {
  if (!bytes) return OXR_OK;
  size_t n16 = 0, head = 0;
  uintptr_t a = (uintptr_t)p;
  if ((a & 15) == 0) n16 = bytes / 16;
  head = n16 * 16;
  size_t ntail = bytes - head;
  size_t threads = 256;
  size_t blocks = std::min<size_t>((std::max(n16, ntail) + threads - 1) / threads, 4096);
  if (blocks == 0) blocks = 1;
  if (ntail > blocks * threads) {
    // unaligned large region: byte loop on the host side of the kernel
    blocks = (ntail + threads - 1) / threads;
  }
  hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3((unsigned)threads), 0, s->s,
                     (uint4 *)p, n16, (uint8_t *)p + head, ntail);
  OXR_HIP(hipGetLastError());
  return OXR_OK;
}

StreamCore *retain(StreamCore *c) {
  if (c) c->refs.fetch_add(1, std::memory_order_relaxed);
  return c;
}

void release(StreamCore *c) {
  if (!c) return;
  if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  if (c->owned) {
    hipStreamSynchronize(c->s);
    hipStreamDestroy(c->s);
  }
  delete c;
}

SharedBuffers::~SharedBuffers() {
  for (void *p : ptrs)
    if (p) dev->arena->release(p, core->s);
  release(core);
}

std::shared_ptr<void> share_columns(const oxr_batch *cb) {
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  oxr_batch *b = const_cast<oxr_batch *>(cb);  // ownership moves; contents do not change
  bool owned = false;
  for (const DevCol &c : b->cols) owned |= c.owned;
  if (owned) {
    auto sb = std::make_shared<SharedBuffers>();
    sb->dev = b->dev;
    sb->core = retain(b->owner);
    for (DevCol &c : b->cols) {
      if (!c.owned) continue;
      sb->ptrs.push_back(c.values);
      sb->ptrs.push_back(c.validity);
      c.owned = false;
    }
    b->holders.push_back(sb);
  }
  return std::make_shared<std::vector<std::shared_ptr<void>>>(b->holders);
}

oxr_batch *new_batch(oxr_stream *s, uint64_t rows) {
  oxr_batch *b = new oxr_batch();
  b->dev = s->dev;
  b->owner = retain(s->core);
  b->rows = rows;
  return b;
}

int alloc_column(oxr_stream *s, int32_t type, int32_t dim, uint64_t rows, bool nullable,
                 DevCol *out) {
  DevCol c;
  c.type = type;
  c.dim = dim;
  c.len = rows;
  c.owned = true;
  size_t vbytes = type_width(type, dim) * rows;
  vbytes = (vbytes + 255) / 256 * 256;
  if (vbytes == 0) vbytes = 256;
  OXR_TRY(stream_alloc(s, vbytes, &c.values));
  if (nullable) {
    int rc = stream_alloc(s, validity_alloc(rows) + 256, (void **)&c.validity);
    if (rc != OXR_OK) {
      stream_release(s, c.values);
      return rc;
    }
    rc = fill_zero(s, c.validity, validity_alloc(rows));
    if (rc != OXR_OK) return rc;
  }
  *out = c;
  return OXR_OK;
}

// -------------------------------------------------------------------- tuning
namespace {
std::mutex g_tune_mu;
std::map<std::string, int64_t> g_tune;

// Every knob the library reads, with the values its kernels take. A value
// outside them is refused (OXR_EPLAN), so a benchmark never reports a shape
// that did not run.
struct KnobSpec {
  const char *name;
  int64_t lo, hi;      // inclusive range
  uint32_t allowed;    // nonzero: bit v set for each allowed value v < 32
};
constexpr uint32_t kBool = 0x3u;
const KnobSpec kKnobs[] = {
    {"filter.jit", 0, 1, kBool},
    {"filter.rows", 1, 32, 0},               // generated filter: rows per lane
    {"filter.persistent", 0, 1, kBool},      // generated filter: persistent grid, tiles by ticket
    {"filter.grid", 1, 16, 0},               // persistent filter: workgroups per CU (default: what fits)
    {"filter.waves", 4, 16, (1u << 4) | (1u << 8) | (1u << 12) | (1u << 16)},  // generated filter: waves per workgroup
    {"filter.wg", 0, 2, 0},                  // one workgroup per batch: 0 never (default), 1 many batches, 2 always
    {"filter.staged", 0, 2, 0},              // generated filter, kept rows staged in LDS: 0 never, 1 by the shape's kept fraction (default), 2 always
    {"filter.staged_waves", 4, 16, (1u << 4) | (1u << 8) | (1u << 12) | (1u << 16)},  // its waves per workgroup (default 8)
    {"filter.chunk_rows", 1, 16, 0},         // its rows per lane per chunk (default 4; times waves at most 64)
    {"filter.chunks", 1, 32, 0},             // its chunks per tile (default 8; times chunk rows at most 32)
    {"filter.stage_rows", 8, 8192, 0},       // its LDS stage in kept rows (default 3072; rounded down to 8)
    {"filter.stage_nt", 0, 1, kBool},        // its copy out of the stage with nontemporal stores (default 1)
    {"agg.spill_nt", 0, 1, kBool},           // aggregation spill: scatter and fold read records with nontemporal loads (default 1)
    {"filter.wg_rows", 1, 32, 0},            // its rows per lane per register set (default 8)
    {"filter.wg_grid", 1, 8, 0},             // its workgroups per CU (default 1)
    {"filter.wg_waves", 4, 16, (1u << 4) | (1u << 8) | (1u << 16)},  // its waves per workgroup (default 8)
    {"filter.v", 1, 4, (1u << 1) | (1u << 2) | (1u << 4)},
    {"filter.g", 1, 4, (1u << 1) | (1u << 2) | (1u << 4)},
    {"jit.params", 0, 1, kBool},
    {"jit.stream_loads", 0, 1, kBool},       // generated kernels: nontemporal column loads
    {"jit.stream_stores", 0, 1, kBool},      // generated filter: nontemporal projection stores
    {"transfer.staging", 0, 1, kBool},       // upload / download through the stream's pinned ring
    {"transfer.chunk_kib", 64, 65536, 0},    // its chunk size in KiB (default 1024)
    {"agg.jit", 0, 1, kBool},
    {"agg.rows", 1, 32, 0},                  // generated aggregate: rows per lane
    {"agg.grid", 1, 64, 0},
    {"agg.jit_grid", 1, 64, 0},
    {"agg.exact_f64", 0, 1, kBool},
    {"agg.dense", 0, 1, kBool},
    {"agg.range_rows", 0, (int64_t)1 << 48, 0},
    {"agg.range_cache", 0, 1, kBool},          // key ranges kept with the batch (default 1)
    {"agg.hot_cache", 0, 1, kBool},            // hot-group maps kept with the batch (default 1)
    {"agg.spill", 0, 1, kBool},
    {"agg.spill_rows", 0, (int64_t)1 << 48, 0},
    {"agg.spill_pct", 1, 400, 0},
    {"agg.hot", 0, 1, kBool},                 // hot-group kernel for wide skewed key ranges
    {"agg.hot_rows", 1, (int64_t)1 << 48, 0},  // rows from which an update samples for hot groups
    {"agg.hot_lane_rows", 1, 32, 0},           // hot-group kernel: rows per lane (default: the shape's)
    {"agg.hot_grid", 1, 64, 0},                // hot-group kernel: workgroups per CU (default 8)
    {"agg.keyless", 0, 1, kBool},              // global aggregates: the keyless kernel (default 1)
    {"agg.keyless_grid", 1, 64, 0},            // keyless kernel: workgroups per CU (default 4)
    {"agg.multi_rows", 1, 16, 0},              // batched requests: rows per lane (default 4)
    {"agg.multi_grid", 1, 16, 0},              // batched requests: workgroups per CU (default 2)
    {"agg.batch_graph", 0, 1, kBool},          // batched requests: replay a captured hipGraph (default 1)
    {"vector.resident_queries", 0, 1, kBool},
    {"vector.gemm", 0, 1, kBool},              // batches of 129-256 queries: the GEMM-tiled scan
    {"vector.wide", 0, 3, 0},
    {"vector.nt8", 0, 1, kBool},
    {"vector.pilot", 0, 1, kBool},
    {"vector.select_sampled", 0, 4096, 0},    // selection: sampled bound then one sort for C at least this (default 128; 0 off)
    {"vector.pilot_rows", 1, (int64_t)1 << 40, 0},
    {"vector.pilot_cap", 1, (int64_t)1 << 32, 0},
};
}  // namespace

// OXR_OK when `knob` is known and takes `value` (a negative value resets a
// known knob to its default).
int tune_check(const char *knob, int64_t value) {
  for (const KnobSpec &k : kKnobs) {
    if (std::strcmp(k.name, knob) != 0) continue;
    if (value < 0) return OXR_OK;
    const bool in_range = value >= k.lo && value <= k.hi;
    if (in_range && (!k.allowed || (value < 32 && ((k.allowed >> value) & 1u)))) return OXR_OK;
    return fail(OXR_EPLAN, "rocm.tune",
                std::string("knob ") + knob + " does not take the value " + std::to_string(value));
  }
  return fail(OXR_EPLAN, "rocm.tune", std::string("unknown knob ") + knob);
}

std::atomic<uint64_t> g_tune_epoch{0};
uint64_t tune_epoch() { return g_tune_epoch.load(); }

int64_t tune_get(const char *knob, int64_t dflt) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find(knob);
  return it == g_tune.end() ? dflt : it->second;
}

// ----------------------------------------------------------------- profiling
namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof;
std::atomic<bool> g_prof_on{false};
}  // namespace

ProfScope::ProfScope(hipStream_t s, const char *kernel) : s_(s), name_(kernel) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return;
  if (hipEventCreate(&a_) != hipSuccess) {
    a_ = nullptr;
    return;
  }
  hipEventRecord(a_, s_);
}

ProfScope::~ProfScope() {
  if (!a_) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a_);
    return;
  }
  hipEventRecord(b, s_);
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof.push_back({name_, a_, b});
}

}  // namespace oxr

using namespace oxr;

// ===================================================================== C ABI
extern "C" {

const char *oxr_last_error(void) { return t_error.c_str(); }
const char *oxr_last_feature(void) { return t_feature.c_str(); }

int oxr_probe(int ordinal, char *why, size_t why_len) {
  auto say = [&](const std::string &m) {
    if (why && why_len) {
      std::snprintf(why, why_len, "%s", m.c_str());
    }
  };
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    say(std::string("no HIP runtime or driver: ") + hipGetErrorString(e));
    return fail(OXR_EUNSUPPORTED, "rocm.probe", "no HIP device");
  }
  if (ordinal < 0 || ordinal >= n) {
    say("device ordinal " + std::to_string(ordinal) + " out of range (" + std::to_string(n) +
        " devices)");
    return fail(OXR_EUNSUPPORTED, "rocm.probe", "device ordinal out of range");
  }
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, ordinal);
  if (e != hipSuccess) {
    say(std::string("hipGetDeviceProperties failed: ") + hipGetErrorString(e));
    return fail(OXR_EUNSUPPORTED, "rocm.probe", "device properties unavailable");
  }
  std::string arch = p.gcnArchName;
  if (arch.rfind("gfx950", 0) != 0) {
    say("device " + std::to_string(ordinal) + " is " + arch + "; this build targets gfx950 only");
    return fail(OXR_EUNSUPPORTED, "rocm.probe", "not a gfx950 device");
  }
  say("");
  return OXR_OK;
}

int oxr_open(int ordinal, oxr_device **out) {
  static std::mutex mu;
  static std::map<int, oxr_device *> opened;
  std::lock_guard<std::mutex> g(mu);
  auto it = opened.find(ordinal);
  if (it != opened.end()) {
    *out = it->second;
    return OXR_OK;
  }
  char why[256];
  OXR_TRY(oxr_probe(ordinal, why, sizeof why));
  OXR_HIP(hipSetDevice(ordinal));
  hipDeviceProp_t p;
  OXR_HIP(hipGetDeviceProperties(&p, ordinal));
  oxr_device *d = new oxr_device();
  d->ordinal = ordinal;
  d->cu_count = p.multiProcessorCount;
  d->arch = p.gcnArchName;
  d->lds_per_block = p.sharedMemPerBlock;
  d->arena = new Arena(ordinal);
  opened[ordinal] = d;
  *out = d;
  return OXR_OK;
}

int oxr_mem_info(oxr_device *dev, uint64_t *total_bytes, uint64_t *free_bytes,
                 int *pinned_supported) {
  if (!dev) return fail(OXR_EPLAN, "rocm.device", "null device");
  size_t fr = 0, tot = 0;
  OXR_HIP(hipSetDevice(dev->ordinal));
  OXR_HIP(hipMemGetInfo(&fr, &tot));
  if (total_bytes) *total_bytes = tot;
  if (free_bytes) *free_bytes = fr;
  if (pinned_supported) *pinned_supported = 1;
  return OXR_OK;
}

int oxr_arena_info(oxr_device *dev, uint64_t *in_use, uint64_t *cached) {
  if (!dev) return fail(OXR_EPLAN, "rocm.device", "null device");
  dev->arena->stats(in_use, cached);
  return OXR_OK;
}

int oxr_alloc_device(oxr_device *dev, size_t bytes, oxr_buffer **out) {
  if (!dev || !out) return fail(OXR_EPLAN, "rocm.device", "null argument");
  void *p = nullptr;
  OXR_TRY(dev->arena->alloc(bytes ? bytes : 1, nullptr, &p));
  hipError_t e = hipMemsetAsync(p, 0, bytes ? bytes : 1, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    dev->arena->release(p, nullptr);
    return fail_hip(e, "hipMemset");
  }
  oxr_buffer *b = new oxr_buffer();
  b->dev = dev;
  b->ptr = p;
  b->len = bytes;
  *out = b;
  return OXR_OK;
}

int oxr_alloc_pinned(oxr_device *dev, size_t bytes, oxr_buffer **out) {
  if (!dev || !out) return fail(OXR_EPLAN, "rocm.device", "null argument");
  void *p = nullptr;
  OXR_HIP(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
  std::memset(p, 0, bytes ? bytes : 1);
  oxr_buffer *b = new oxr_buffer();
  b->dev = dev;
  b->ptr = p;
  b->len = bytes;
  b->pinned = true;
  *out = b;
  return OXR_OK;
}

void *oxr_buffer_ptr(const oxr_buffer *buf) { return buf ? buf->ptr : nullptr; }
size_t oxr_buffer_len(const oxr_buffer *buf) { return buf ? buf->len : 0; }

void oxr_buffer_free(oxr_buffer *buf) {
  if (!buf) return;
  // a buffer carries no stream: copies into or out of it may still be in
  // flight on any stream, so its memory goes back only after the device is
  // idle (buffers are the trait's raw memory, not the operators' path)
  (void)hipDeviceSynchronize();
  if (buf->pinned) hipHostFree(buf->ptr);
  else buf->dev->arena->release(buf->ptr, nullptr);
  delete buf;
}

extern "C++" {
namespace oxr {
int stream_pinned(oxr_stream *s, size_t bytes, void **out) {
  if (s->pinned_len < bytes) {
    if (s->pinned) (void)hipHostFree(s->pinned);
    s->pinned = nullptr;
    s->pinned_len = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    OXR_HIP(hipHostMalloc(&s->pinned, want, hipHostMallocDefault));
    s->pinned_len = want;
  }
  *out = s->pinned;
  return OXR_OK;
}
}  // namespace oxr
}

int oxr_stream_create(oxr_device *dev, oxr_stream **out) {
  if (!dev || !out) return fail(OXR_EPLAN, "rocm.stream", "null argument");
  OXR_HIP(hipSetDevice(dev->ordinal));
  hipStream_t s;
  OXR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  oxr_stream *st = new oxr_stream();
  st->dev = dev;
  st->core = new StreamCore();
  st->core->s = s;
  st->core->owned = true;
  st->s = s;
  *out = st;
  return OXR_OK;
}

int oxr_stream_wrap(oxr_device *dev, void *hip_stream, oxr_stream **out) {
  if (!dev || !out) return fail(OXR_EPLAN, "rocm.stream", "null argument");
  oxr_stream *st = new oxr_stream();
  st->dev = dev;
  st->core = new StreamCore();
  st->core->s = (hipStream_t)hip_stream;
  st->core->owned = false;
  st->s = st->core->s;
  *out = st;
  return OXR_OK;
}

void oxr_stream_destroy(oxr_stream *s) {
  if (!s) return;
  if (s->pinned) (void)hipHostFree(s->pinned);
  // no copy through the ring is in flight: upload and download return
  // after their stream synchronised
  for (int i = 0; i < 2; i++) {
    if (s->stage[i]) (void)hipHostFree(s->stage[i]);
    if (s->stage_ev[i]) (void)hipEventDestroy(s->stage_ev[i]);
  }
  // the HIP stream itself goes when the last batch/table allocated on it does
  release(s->core);
  delete s;
}

int oxr_stream_sync(oxr_stream *s) {
  if (!s) return fail(OXR_EPLAN, "rocm.stream", "null stream");
  OXR_HIP(hipStreamSynchronize(s->s));
  return OXR_OK;
}

int oxr_copy_h2d(oxr_stream *s, const void *src, oxr_buffer *dst, size_t bytes) {
  if (!s || !dst) return fail(OXR_EPLAN, "rocm.copy", "null argument");
  if (bytes > dst->len) return fail(OXR_EEXECUTION, "rocm.copy", "copy exceeds the destination");
  OXR_HIP(hipMemcpyAsync(dst->ptr, src, bytes, hipMemcpyHostToDevice, s->s));
  return OXR_OK;
}

int oxr_copy_d2h(oxr_stream *s, const oxr_buffer *src, void *dst, size_t bytes) {
  if (!s || !src) return fail(OXR_EPLAN, "rocm.copy", "null argument");
  if (bytes > src->len) return fail(OXR_EEXECUTION, "rocm.copy", "copy exceeds the source");
  OXR_HIP(hipMemcpyAsync(dst, src->ptr, bytes, hipMemcpyDeviceToHost, s->s));
  return OXR_OK;
}

// ------------------------------------------------------------------ batches

static int arrow_type(const struct ArrowSchema *f, int32_t *type, int32_t *dim) {
  const char *fmt = f->format ? f->format : "";
  *dim = 0;
  if (!std::strcmp(fmt, "i")) *type = OXR_INT32;
  else if (!std::strcmp(fmt, "l")) *type = OXR_INT64;
  else if (!std::strcmp(fmt, "f")) *type = OXR_FLOAT32;
  else if (!std::strcmp(fmt, "g")) *type = OXR_FLOAT64;
  else if (!std::strncmp(fmt, "+w:", 3) && f->n_children == 1 && f->children[0]->format &&
           !std::strcmp(f->children[0]->format, "f")) {
    *type = OXR_F32_VECTOR;
    *dim = std::atoi(fmt + 3);
    if (*dim <= 0)
      return fail(OXR_EUNSUPPORTED, "device.column_type", "FixedSizeList of size 0");
  } else {
    return fail(OXR_EUNSUPPORTED, "device.column_type",
                std::string("Arrow format '") + fmt + "' columns are not device-resident");
  }
  return OXR_OK;
}

static inline bool bit_at(const uint8_t *bits, uint64_t i) { return (bits[i >> 3] >> (i & 7)) & 1; }

// The validity bitmap starting at bit 0 (columns.rs:74-78: re-packed only
// when the slice is not byte-aligned).
static std::vector<uint8_t> repack_bits(const uint8_t *bits, uint64_t offset, uint64_t len) {
  std::vector<uint8_t> out(validity_alloc(len), 0);
  if ((offset & 7) == 0) {
    std::memcpy(out.data(), bits + offset / 8, validity_bytes(len));
    if (len & 7) out[len / 8] &= (uint8_t)((1u << (len & 7)) - 1);
    return out;
  }
  for (uint64_t i = 0; i < len; i++)
    if (bit_at(bits, offset + i)) out[i >> 3] |= (uint8_t)(1u << (i & 7));
  return out;
}

// ------------------------------------------------------ pinned staging ring
// Arrow buffers arrive in pageable host memory. With the knob
// transfer.staging on (the default), upload and download go through the
// stream's two pinned halves of transfer.chunk_kib KiB: the host copy of
// one chunk runs while the other half's DMA is in flight, and a half is
// reused only once its last DMA has finished (its event). Off, every buffer
// is one hipMemcpyAsync from pageable memory.
namespace {
struct HostCopy {
  void *dst;
  const void *src;
  size_t bytes;
};

int stage_ring(oxr_stream *s, size_t *chunk) {
  const size_t want = (size_t)tune_get("transfer.chunk_kib", 1024) * 1024;
  if (s->stage_len != want) {
    for (int i = 0; i < 2; i++) {
      if (s->stage[i]) (void)hipHostFree(s->stage[i]);
      s->stage[i] = nullptr;
    }
    s->stage_len = 0;
    for (int i = 0; i < 2; i++) {
      OXR_HIP(hipHostMalloc(&s->stage[i], want, hipHostMallocDefault));
      if (!s->stage_ev[i]) OXR_HIP(hipEventCreateWithFlags(&s->stage_ev[i], hipEventDisableTiming));
    }
    s->stage_len = want;
  }
  *chunk = want;
  return OXR_OK;
}

// Host to device through the ring; returns with the last DMAs queued.
int stage_h2d(oxr_stream *s, const std::vector<HostCopy> &copies) {
  size_t chunk = 0;
  OXR_TRY(stage_ring(s, &chunk));
  bool busy[2] = {false, false};
  int slot = 0;
  for (const HostCopy &c : copies) {
    for (size_t off = 0; off < c.bytes; off += chunk) {
      const size_t m = std::min(chunk, c.bytes - off);
      if (busy[slot]) OXR_HIP(hipEventSynchronize(s->stage_ev[slot]));
      std::memcpy(s->stage[slot], (const char *)c.src + off, m);
      OXR_HIP(hipMemcpyAsync((char *)c.dst + off, s->stage[slot], m, hipMemcpyHostToDevice, s->s));
      OXR_HIP(hipEventRecord(s->stage_ev[slot], s->s));
      busy[slot] = true;
      slot ^= 1;
    }
  }
  return OXR_OK;
}

// Device to host through the ring: two DMAs in flight, each chunk copied
// out of its half as soon as its DMA is done; returns with every byte home.
int stage_d2h(oxr_stream *s, const std::vector<HostCopy> &copies) {
  size_t chunk = 0;
  OXR_TRY(stage_ring(s, &chunk));
  struct Pending {
    int slot;
    void *dst;
    size_t bytes;
  };
  Pending pend[2];
  int npend = 0, head = 0, slot = 0;
  auto drain_one = [&]() -> int {
    const Pending p = pend[head];
    head ^= 1;
    npend--;
    OXR_HIP(hipEventSynchronize(s->stage_ev[p.slot]));
    std::memcpy(p.dst, s->stage[p.slot], p.bytes);
    return OXR_OK;
  };
  for (const HostCopy &c : copies) {
    for (size_t off = 0; off < c.bytes; off += chunk) {
      const size_t m = std::min(chunk, c.bytes - off);
      if (npend == 2) OXR_TRY(drain_one());  // the oldest chunk used `slot`
      OXR_HIP(hipMemcpyAsync(s->stage[slot], (const char *)c.src + off, m, hipMemcpyDeviceToHost, s->s));
      OXR_HIP(hipEventRecord(s->stage_ev[slot], s->s));
      pend[(head + npend) & 1] = {slot, (char *)c.dst + off, m};
      npend++;
      slot ^= 1;
    }
  }
  while (npend) OXR_TRY(drain_one());
  return OXR_OK;
}

bool staging_on() { return tune_get("transfer.staging", 1) > 0; }
}  // namespace

int oxr_upload(oxr_stream *s, const struct ArrowSchema *schema, const struct ArrowArray *array,
               oxr_batch **out) {
  clear_error();
  if (!s || !schema || !array || !out) return fail(OXR_EPLAN, "rocm.upload", "null argument");
  if (std::strcmp(schema->format ? schema->format : "", "+s") != 0)
    return fail(OXR_EFORMAT, "rocm.upload", "upload expects an Arrow struct array (a RecordBatch)");
  if (schema->n_children != array->n_children)
    return fail(OXR_EFORMAT, "rocm.upload", "schema and array disagree on the column count");
  const uint64_t rows = (uint64_t)array->length;
  const uint64_t poff = (uint64_t)array->offset;
  // check every column's type first: an ineligible column uploads nothing
  std::vector<int32_t> types(schema->n_children), dims(schema->n_children);
  for (int64_t i = 0; i < schema->n_children; i++)
    OXR_TRY(arrow_type(schema->children[i], &types[i], &dims[i]));
  std::unique_ptr<oxr_batch> b(new_batch(s, rows));
  auto cleanup = [&]() {
    for (auto &c : b->cols) {
      if (c.owned) {
        stream_release(s, c.values);
        stream_release(s, c.validity);
      }
    }
  };
  std::vector<std::vector<uint8_t>> staged;  // kept alive until the copies finish
  staged.reserve(schema->n_children);        // (their addresses go into `copies`)
  std::vector<HostCopy> copies;
  for (int64_t i = 0; i < schema->n_children; i++) {
    const struct ArrowArray *a = array->children[i];
    const struct ArrowSchema *f = schema->children[i];
    const uint64_t off = poff + (uint64_t)a->offset;
    const uint8_t *valid = a->n_buffers > 0 ? (const uint8_t *)a->buffers[0] : nullptr;
    const bool nullable = valid != nullptr && a->null_count != 0;
    const void *values = nullptr;
    uint64_t child_off = 0;
    if (types[i] == OXR_F32_VECTOR) {
      const struct ArrowArray *ch = a->children[0];
      values = ch->n_buffers > 1 ? ch->buffers[1] : nullptr;
      child_off = (uint64_t)ch->offset;
      // columns.rs:118-187: a NULL element inside a valid row cannot travel
      const uint8_t *ev = ch->n_buffers > 0 ? (const uint8_t *)ch->buffers[0] : nullptr;
      if (ev && ch->null_count != 0) {
        const uint64_t d = (uint64_t)dims[i];
        for (uint64_t r = 0; r < rows; r++) {
          if (nullable && !bit_at(valid, off + r)) continue;
          for (uint64_t e = 0; e < d; e++) {
            if (!bit_at(ev, child_off + (off + r) * d + e)) {
              cleanup();
              return fail(OXR_EUNSUPPORTED, "device.vector_null_element",
                          "row " + std::to_string(r) + " of vector column '" +
                              (f->name ? f->name : "") +
                              "' is valid but holds a NULL element, which the device format "
                              "(one validity bit per row) would return as a number");
            }
          }
        }
      }
    } else {
      values = a->n_buffers > 1 ? a->buffers[1] : nullptr;
    }
    DevCol c;
    int rc = alloc_column(s, types[i], dims[i], rows, nullable, &c);
    if (rc != OXR_OK) {
      cleanup();
      return rc;
    }
    b->cols.push_back(c);
    b->names.push_back(f->name ? f->name : "");
    const size_t w = type_width(types[i], dims[i]);
    if (rows) {
      if (!values) {
        cleanup();
        return fail(OXR_EFORMAT, "rocm.upload", "column without a value buffer");
      }
      const uint8_t *src = (const uint8_t *)values;
      if (types[i] == OXR_F32_VECTOR) src += (child_off + off * (uint64_t)dims[i]) * 4;
      else src += off * w;
      copies.push_back({c.values, src, w * rows});
      if (nullable) {
        staged.push_back(repack_bits(valid, off, rows));
        copies.push_back({c.validity, staged.back().data(), staged.back().size()});
      }
    }
  }
  if (staging_on()) {
    const int rc = stage_h2d(s, copies);
    if (rc != OXR_OK) {
      (void)hipStreamSynchronize(s->s);  // nothing still reads the host buffers
      cleanup();
      return rc;
    }
  } else {
    for (const HostCopy &c : copies) {
      const hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, s->s);
      if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->s);
        cleanup();
        return fail_hip(e, "hipMemcpyAsync(upload)");
      }
    }
  }
  hipError_t e = hipStreamSynchronize(s->s);
  if (e != hipSuccess) {
    cleanup();
    return fail_hip(e, "hipStreamSynchronize(upload)");
  }
  *out = b.release();
  return OXR_OK;
}

// ----------------------------------------------------------- Arrow export
namespace {
struct ExportedArray {
  std::vector<void *> mem;                  // host allocations to free
  std::vector<const void *> buffers;        // this array's buffers
  std::vector<struct ArrowArray *> children;
};
struct ExportedSchema {
  std::string format, name;
  std::vector<struct ArrowSchema *> children;
};

void release_array(struct ArrowArray *a) {
  if (!a || !a->release) return;
  for (int64_t i = 0; i < a->n_children; i++) {
    struct ArrowArray *c = a->children[i];
    if (c && c->release) c->release(c);
    delete c;
  }
  ExportedArray *p = (ExportedArray *)a->private_data;
  for (void *m : p->mem) std::free(m);
  delete p;
  a->release = nullptr;
}

void release_schema(struct ArrowSchema *s) {
  if (!s || !s->release) return;
  for (int64_t i = 0; i < s->n_children; i++) {
    struct ArrowSchema *c = s->children[i];
    if (c && c->release) c->release(c);
    delete c;
  }
  delete (ExportedSchema *)s->private_data;
  s->release = nullptr;
}

struct ArrowSchema *make_schema(const std::string &format, const std::string &name, bool nullable) {
  struct ArrowSchema *s = new struct ArrowSchema();
  ExportedSchema *p = new ExportedSchema();
  p->format = format;
  p->name = name;
  s->format = p->format.c_str();
  s->name = p->name.c_str();
  s->metadata = nullptr;
  s->flags = nullable ? ARROW_FLAG_NULLABLE : 0;
  s->n_children = 0;
  s->children = nullptr;
  s->dictionary = nullptr;
  s->release = release_schema;
  s->private_data = p;
  return s;
}

void *host_alloc(size_t bytes) {
  void *p = nullptr;
  if (posix_memalign(&p, 64, bytes ? (bytes + 63) / 64 * 64 : 64) != 0) return nullptr;
  return p;
}

int64_t count_nulls(const uint8_t *bits, uint64_t len) {
  int64_t valid = 0;
  uint64_t full = len / 8;
  for (uint64_t i = 0; i < full; i++) valid += __builtin_popcount(bits[i]);
  for (uint64_t i = full * 8; i < len; i++) valid += bit_at(bits, i);
  return (int64_t)len - valid;
}
}  // namespace

int oxr_download(oxr_stream *s, const oxr_batch *batch, struct ArrowSchema *out_schema,
                 struct ArrowArray *out_array) {
  clear_error();
  if (!s || !batch || !out_schema || !out_array)
    return fail(OXR_EPLAN, "rocm.download", "null argument");
  const uint64_t rows = batch->rows;
  const size_t ncols = batch->cols.size();
  // top-level struct
  ExportedArray *top = new ExportedArray();
  ExportedSchema *tops = new ExportedSchema();
  tops->format = "+s";
  tops->name = "";
  std::vector<struct ArrowArray *> arrays;
  std::vector<struct ArrowSchema *> schemas;
  std::vector<std::pair<uint8_t *, uint64_t>> validity_to_count;
  auto abort_all = [&](int rc) {
    for (auto *a : arrays) {
      release_array(a);
      delete a;
    }
    for (auto *sc : schemas) {
      release_schema(sc);
      delete sc;
    }
    for (void *m : top->mem) std::free(m);
    delete top;
    delete tops;
    return rc;
  };
  std::vector<struct ArrowArray *> nulls_fix;
  std::vector<HostCopy> copies;
  for (size_t i = 0; i < ncols; i++) {
    const DevCol &c = batch->cols[i];
    const size_t w = type_width(c.type, c.dim);
    std::string fmt;
    switch (c.type) {
      case OXR_INT32: fmt = "i"; break;
      case OXR_INT64: fmt = "l"; break;
      case OXR_FLOAT32: fmt = "f"; break;
      case OXR_FLOAT64: fmt = "g"; break;
      case OXR_F32_VECTOR: fmt = "+w:" + std::to_string(c.dim); break;
    }
    const std::string name = i < batch->names.size() ? batch->names[i] : "c" + std::to_string(i);
    struct ArrowSchema *sc = make_schema(fmt, name, true);
    schemas.push_back(sc);
    ExportedArray *p = new ExportedArray();
    struct ArrowArray *a = new struct ArrowArray();
    std::memset(a, 0, sizeof *a);
    a->private_data = p;
    a->release = release_array;
    arrays.push_back(a);
    a->length = (int64_t)rows;
    a->offset = 0;
    uint8_t *vbits = nullptr;
    if (c.validity && rows) {
      vbits = (uint8_t *)host_alloc(validity_bytes(rows));
      if (!vbits) return abort_all(fail(OXR_EALLOCATION, "rocm.download", "host allocation"));
      p->mem.push_back(vbits);
      copies.push_back({vbits, c.validity, validity_bytes(rows)});
    }
    void *vals = host_alloc(w * rows);
    if (!vals) return abort_all(fail(OXR_EALLOCATION, "rocm.download", "host allocation"));
    if (rows) copies.push_back({vals, c.values, w * rows});
    if (c.type == OXR_F32_VECTOR) {
      // parent: validity only; child: Float32 values
      p->buffers = {vbits};
      a->n_buffers = 1;
      a->buffers = p->buffers.data();
      ExportedArray *cp = new ExportedArray();
      struct ArrowArray *ch = new struct ArrowArray();
      std::memset(ch, 0, sizeof *ch);
      cp->mem.push_back(vals);
      cp->buffers = {nullptr, vals};
      ch->length = (int64_t)(rows * (uint64_t)c.dim);
      ch->null_count = 0;
      ch->n_buffers = 2;
      ch->buffers = cp->buffers.data();
      ch->private_data = cp;
      ch->release = release_array;
      p->children.push_back(ch);
      a->n_children = 1;
      a->children = p->children.data();
      ExportedSchema *sp = (ExportedSchema *)sc->private_data;
      sp->children.push_back(make_schema("f", "item", true));
      sc->n_children = 1;
      sc->children = sp->children.data();
    } else {
      p->mem.push_back(vals);
      p->buffers = {vbits, vals};
      a->n_buffers = 2;
      a->buffers = p->buffers.data();
    }
    if (vbits) {
      validity_to_count.push_back({vbits, rows});
      nulls_fix.push_back(a);
    } else {
      a->null_count = 0;
    }
  }
  if (staging_on()) {
    const int rc = stage_d2h(s, copies);
    if (rc != OXR_OK) {
      (void)hipStreamSynchronize(s->s);  // nothing still writes the host buffers
      return abort_all(rc);
    }
  } else {
    for (const HostCopy &c : copies) {
      const hipError_t e = hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s->s);
      if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->s);
        return abort_all(fail_hip(e, "hipMemcpyAsync(download)"));
      }
    }
  }
  hipError_t e = hipStreamSynchronize(s->s);
  if (e != hipSuccess) return abort_all(fail_hip(e, "hipStreamSynchronize(download)"));
  for (size_t k = 0; k < nulls_fix.size(); k++)
    nulls_fix[k]->null_count = count_nulls(validity_to_count[k].first, validity_to_count[k].second);
  // assemble the struct
  top->children = arrays;
  top->buffers = {nullptr};
  std::memset(out_array, 0, sizeof *out_array);
  out_array->length = (int64_t)rows;
  out_array->null_count = 0;
  out_array->offset = 0;
  out_array->n_buffers = 1;
  out_array->buffers = top->buffers.data();
  out_array->n_children = (int64_t)ncols;
  out_array->children = top->children.data();
  out_array->private_data = top;
  out_array->release = release_array;
  tops->children = schemas;
  std::memset(out_schema, 0, sizeof *out_schema);
  out_schema->format = tops->format.c_str();
  out_schema->name = tops->name.c_str();
  out_schema->flags = 0;
  out_schema->n_children = (int64_t)ncols;
  out_schema->children = tops->children.data();
  out_schema->private_data = tops;
  out_schema->release = release_schema;
  return OXR_OK;
}

int oxr_batch_wrap(oxr_device *dev, const oxr_column *columns, const char *const *names,
                   int32_t ncols, uint64_t rows, oxr_batch **out) {
  clear_error();
  if (!dev || !out || (ncols > 0 && !columns)) return fail(OXR_EPLAN, "rocm.batch", "null argument");
  oxr_batch *b = new oxr_batch();
  b->dev = dev;
  b->rows = rows;
  for (int32_t i = 0; i < ncols; i++) {
    const oxr_column &c = columns[i];
    if (c.type < OXR_INT32 || c.type > OXR_F32_VECTOR || (c.type == OXR_F32_VECTOR && c.dim <= 0)) {
      delete b;
      return fail(OXR_EUNSUPPORTED, "device.column_type", "unknown column type");
    }
    if (c.len != rows) {
      delete b;
      return fail(OXR_EFORMAT, "rocm.batch", "column length differs from the batch's rows");
    }
    if (rows && !c.values) {
      delete b;
      return fail(OXR_EFORMAT, "rocm.batch", "column without values");
    }
    DevCol d;
    d.type = c.type;
    d.dim = c.type == OXR_F32_VECTOR ? c.dim : 0;
    d.len = rows;
    d.values = c.values;
    d.validity = c.validity;
    d.owned = false;
    b->cols.push_back(d);
    b->names.push_back(names && names[i] ? names[i] : "c" + std::to_string(i));
  }
  *out = b;
  return OXR_OK;
}

uint64_t oxr_batch_rows(const oxr_batch *b) { return b ? b->rows : 0; }
int32_t oxr_batch_ncols(const oxr_batch *b) { return b ? (int32_t)b->cols.size() : 0; }

int oxr_batch_column(const oxr_batch *b, int32_t i, oxr_column *out) {
  if (!b || !out) return fail(OXR_EPLAN, "rocm.batch", "null argument");
  if (i < 0 || i >= (int32_t)b->cols.size())
    return fail(OXR_EPLAN, "rocm.batch", "column index out of range");
  const DevCol &c = b->cols[i];
  out->type = c.type;
  out->dim = c.dim;
  out->len = c.len;
  out->values = c.values;
  out->validity = c.validity;
  return OXR_OK;
}

const char *oxr_batch_name(const oxr_batch *b, int32_t i) {
  if (!b || i < 0 || i >= (int32_t)b->names.size()) return "";
  return b->names[i].c_str();
}

int oxr_tune_set(const char *knob, int64_t value) {
  if (!knob) return fail(OXR_EPLAN, "rocm.tune", "null knob");
  OXR_TRY(tune_check(knob, value));
  std::lock_guard<std::mutex> g(g_tune_mu);
  if (value < 0) g_tune.erase(knob);
  else g_tune[knob] = value;
  g_tune_epoch.fetch_add(1);
  return OXR_OK;
}

int oxr_tune_get(const char *knob, int64_t *value) {
  if (!knob || !value) return fail(OXR_EPLAN, "rocm.tune", "null argument");
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find(knob);
  *value = it == g_tune.end() ? -1 : it->second;
  return OXR_OK;
}

void oxr_profile_enable(int on) { g_prof_on.store(on != 0); }

void oxr_profile_reset(void) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto &r : g_prof) {
    hipEventSynchronize(r.b);
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  g_prof.clear();
}

int oxr_profile_get(const char *kernel, double *total_ms, uint64_t *launches) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  double tot = 0;
  uint64_t n = 0;
  for (auto &r : g_prof) {
    if (kernel && r.name != kernel) continue;
    OXR_HIP(hipEventSynchronize(r.b));
    float ms = 0;
    OXR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    n++;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = n;
  return OXR_OK;
}

void oxr_batch_free(oxr_batch *b) {
  if (!b) return;
  hipStream_t s = b->owner ? b->owner->s : nullptr;
  for (auto &c : b->cols) {
    if (!c.owned) continue;
    b->dev->arena->release(c.values, s);
    b->dev->arena->release(c.validity, s);
  }
  for (auto &h : b->hot_maps) {
    b->dev->arena->release(h.mem, h.core->s);
    if (h.ready) (void)hipEventDestroy(h.ready);
    release(h.core);
  }
  release(b->owner);
  delete b;
}

void oxr_batch_free_many(oxr_batch *const *batches, uint64_t n) {
  if (!batches) return;
  for (uint64_t i = 0; i < n; i++) oxr_batch_free(batches[i]);
}

}  // extern "C"

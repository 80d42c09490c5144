// Internal declarations shared by the HIP translation units of liboxide_rocm.
// Not part of the C ABI (include/oxide_rocm.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oxide_rocm.h"

namespace oxr {

// ------------------------------------------------------------------ errors
// Thread-local last error, mirroring EngineError's variants
// (crates/oxidelake-core/src/error.rs:17-73).
int fail(int code, const char *feature, const std::string &msg);
int fail_hip(hipError_t e, const char *what);
void clear_error();

#define OXR_HIP(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return ::oxr::fail_hip(e_, #call); \
  } while (0)

#define OXR_TRY(call)            \
  do {                           \
    int rc_ = (call);            \
    if (rc_ != OXR_OK) return rc_; \
  } while (0)

// ------------------------------------------------------------------- arena
// HBM arena: a caching allocator over hipMalloc. Freed blocks are kept per
// size class and handed out again, so the per-batch path never calls
// hipMalloc in steady state (the reference allocates every buffer per
// batch, cuda/ops.rs:98-103). A block freed on one stream and reused on
// another is ordered after the free by an event.
class Arena {
 public:
  explicit Arena(int ordinal) : ordinal_(ordinal) {}
  ~Arena();
  int alloc(size_t bytes, hipStream_t s, void **out);
  void release(void *p, hipStream_t s);
  void stats(uint64_t *in_use, uint64_t *cached);
  void trim();

 private:
  struct Block {
    void *ptr;
    hipStream_t stream;  // last stream that used the block
    hipEvent_t freed;    // recorded on `stream` at release
  };
  static size_t round_size(size_t bytes);
  int ordinal_;
  std::mutex mu_;
  std::multimap<size_t, Block> free_;
  std::map<void *, size_t> live_;
  uint64_t in_use_ = 0, cached_ = 0;
};

}  // namespace oxr

struct oxr_device {
  int ordinal = 0;
  int cu_count = 0;
  int xcd_count = 8;
  size_t lds_per_block = 65536;
  std::string arch;
  oxr::Arena *arena = nullptr;
};

namespace oxr {
// A HIP stream shared by its `oxr_stream` handle and by everything allocated
// on it (batches, join tables, matches, aggregation states): those free
// their memory in the stream's order, so the stream lives until the last of
// them is freed, whenever the caller destroys its handle.
struct StreamCore {
  hipStream_t s = nullptr;
  bool owned = false;
  std::atomic<int> refs{1};
};
StreamCore *retain(StreamCore *c);
void release(StreamCore *c);
}  // namespace oxr

namespace oxr {
// A pinned host buffer of at least `bytes`, owned by the stream and reused
// by its calls (a call on a stream runs on one thread at a time), for small
// device-to-host read-backs: a copy into pageable memory is staged through
// the runtime's own pinned buffer and costs more than the copy.
int stream_pinned(oxr_stream *s, size_t bytes, void **out);
}  // namespace oxr

struct oxr_stream {
  oxr_device *dev = nullptr;
  hipStream_t s = nullptr;  // == core->s
  oxr::StreamCore *core = nullptr;
  void *pinned = nullptr;   // small result read-backs (see stream_pinned)
  size_t pinned_len = 0;
  // the pinned staging ring of upload / download: two halves of stage_len
  // bytes, each with the event of the last copy that used it
  void *stage[2] = {nullptr, nullptr};
  size_t stage_len = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

struct oxr_buffer {
  oxr_device *dev = nullptr;
  void *ptr = nullptr;
  size_t len = 0;
  bool pinned = false;
};

namespace oxr {

// One device column. `owned` columns were allocated from the arena and are
// returned to it when the batch is freed.
struct DevCol {
  int32_t type = 0;
  int32_t dim = 0;
  uint64_t len = 0;
  void *values = nullptr;
  uint8_t *validity = nullptr;
  bool owned = false;
};

size_t type_width(int32_t type, int32_t dim);  // bytes per row
inline size_t validity_bytes(uint64_t rows) { return (size_t)((rows + 7) / 8); }
// Validity buffers are allocated rounded up to whole 32-bit words so kernels
// may read and write them word-wise.
inline size_t validity_alloc(uint64_t rows) { return (size_t)((rows + 31) / 32) * 4; }

oxr_device *device_of(oxr_stream *s);
int stream_alloc(oxr_stream *s, size_t bytes, void **out);
void stream_release(oxr_stream *s, void *p);

// A zero-fill kernel rather than hipMemsetAsync: it is safe inside a
// captured hipGraph.
int fill_zero(oxr_stream *s, void *p, size_t bytes);

// Tuning knobs (oxr_tune_set): launch shapes and table sizes chosen by the
// autotuner; `dflt` when a knob was never set.
int64_t tune_get(const char *knob, int64_t dflt);
// Bumped by every oxr_tune_set: what was derived from the knobs before is stale.
uint64_t tune_epoch();

// Kernel timing (oxr_profile_*): brackets one launch with HIP events when
// profiling is enabled; a no-op otherwise.
class ProfScope {
 public:
  ProfScope(hipStream_t s, const char *kernel);
  ~ProfScope();

 private:
  hipStream_t s_;
  const char *name_;
  hipEvent_t a_ = nullptr;
};

}  // namespace oxr

struct oxr_batch {
  oxr_device *dev = nullptr;
  oxr::StreamCore *owner = nullptr;  // stream the owned columns were allocated on
  uint64_t rows = 0;
  std::vector<oxr::DevCol> cols;
  std::vector<std::string> names;
  // buffers this batch's columns point into but does not own (views of
  // another result, or of an input whose buffers it shares)
  std::vector<std::shared_ptr<void>> holders;
  // Column statistics measured once and kept with the batch (a batch's
  // memory does not change while it lives): the value range of a key
  // column, for the aggregation's table sizing and kernel choice.
  struct KeyStats {
    int32_t col = -1;
    int64_t min = 0;
    uint64_t count = 0;
  };
  mutable std::mutex stats_mu;
  mutable std::vector<KeyStats> key_stats;
  // the aggregation's hot-group map of this batch's key column (through a
  // join table when join_serial != 0): which groups the hot kernel keeps in
  // LDS, from a sample of the key column; device memory from the arena,
  // released on `core`'s stream with the batch
  struct HotMap {
    uint64_t join_serial = 0;
    int32_t key_col = -1, build_group = -1;
    int64_t smin = 0;
    uint64_t srange = 0;
    uint32_t nh = 0;
    void *mem = nullptr;
    int32_t *hotmap = nullptr;
    int64_t *hotkeys = nullptr;
    oxr::StreamCore *core = nullptr;
    hipEvent_t ready = nullptr;  // recorded on core's stream after the map is built
  };
  mutable std::vector<HotMap> hot_maps;
};

namespace oxr {
// Arena buffers freed together, on the stream that allocated them, when the
// last batch (or index) sharing them lets go.
struct SharedBuffers {
  oxr_device *dev = nullptr;
  StreamCore *core = nullptr;
  std::vector<void *> ptrs;
  ~SharedBuffers();
};
// Keeps every buffer b's columns point to alive for as long as the returned
// holder lives: b's owned columns move into a SharedBuffers (b still frees
// nothing twice), so views of b may outlive b.
std::shared_ptr<void> share_columns(const oxr_batch *b);

// A batch whose columns are allocated by the library on stream s.
oxr_batch *new_batch(oxr_stream *s, uint64_t rows);
// Allocates a column of `rows` rows (values, and a zeroed validity bitmap
// when `nullable`). Values are padded to a multiple of 256 bytes.
int alloc_column(oxr_stream *s, int32_t type, int32_t dim, uint64_t rows, bool nullable,
                 DevCol *out);
}  // namespace oxr

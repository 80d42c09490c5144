/*
 * oxide_rocm.h — C ABI of the MI355X (gfx950) backend for OxideLake's four
 * device operators: fused filter + projection, inner hash join (build, probe,
 * gather), grouped aggregation and vector distance (+ batched top-k).
 *
 * Every entry point here replaces one method of the object-safe Rust trait
 * `GpuBackend` (crates/oxidelake-device/src/backend.rs:44-198); the comment on
 * each names the method and line. A Rust `RocmBackend` implements that trait
 * as a thin `unsafe extern "C"` wrapper over these functions.
 *
 * Conventions
 *  - Every function returns an `int` status (OXR_OK or an OXR_E* code). The
 *    codes mirror `EngineError` (crates/oxidelake-core/src/error.rs:17-73).
 *    OXR_EUNSUPPORTED is the fallback signal: "decline this batch, run the CPU
 *    reference" (crates/oxidelake-compute/src/operator.rs:532-545). Any other
 *    error fails the query.
 *  - The message (and for OXR_EUNSUPPORTED the stable feature id) of the last
 *    failure on the calling thread is in oxr_last_error()/oxr_last_feature().
 *  - Handles are opaque. Every *_free/_destroy accepts NULL.
 *  - Columns exchanged with the host use the Arrow C Data Interface: the
 *    reference's transfer format (crates/oxidelake-device/src/columns.rs:25-40)
 *    is exactly Arrow's buffers: values packed from row 0, an LSB-first
 *    validity bitmap from bit 0, FixedSizeList<Float32> child values.
 *  - Entry points are callable concurrently from different host threads on
 *    different streams (the trait is `Send + Sync`, backend.rs:44).
 */
#ifndef OXIDE_ROCM_H
#define OXIDE_ROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  OXR_OK = 0,
  OXR_EUNSUPPORTED = 1, /* EngineError::Unsupported  error.rs:59-64 */
  OXR_EDEVICE = 2,      /* EngineError::Device       error.rs:33-39 */
  OXR_EEXECUTION = 3,   /* EngineError::Execution    error.rs:42 */
  OXR_EPLAN = 4,        /* EngineError::Plan         error.rs:46 */
  OXR_EFORMAT = 5,      /* EngineError::Format       error.rs:50 */
  OXR_EALLOCATION = 6   /* EngineError::Allocation   error.rs:22-30 */
};

/* Message of the calling thread's last failure ("" when none). */
const char *oxr_last_error(void);
/* Stable feature id of the last OXR_EUNSUPPORTED, e.g. "rocm.filter"
 * (`EngineError::Unsupported::feature`, error.rs:61). */
const char *oxr_last_feature(void);

/* ------------------------------------------------------------------ types */
/* Column types that live on the device: the reference's device-eligible set
 * (columns.rs:48-63: Int64, Float64, Float32, FixedSizeList<Float32>) plus
 * Int32, which the benchmark tables' keys and filter columns use. */
enum {
  OXR_INT32 = 1,
  OXR_INT64 = 2,
  OXR_FLOAT32 = 3,
  OXR_FLOAT64 = 4,
  OXR_F32_VECTOR = 5 /* FixedSizeList<Float32, dim> */
};

/* One column in device memory (or, for oxr_batch_column, a view of one). */
typedef struct oxr_column {
  int32_t type;            /* OXR_INT32 .. OXR_F32_VECTOR */
  int32_t dim;             /* list size for OXR_F32_VECTOR, else 0 */
  uint64_t len;            /* rows */
  void *values;            /* len * width bytes, packed from row 0 */
  uint8_t *validity;       /* ceil(len/8) bytes, bit i = row i; NULL = all valid */
} oxr_column;

typedef struct oxr_device oxr_device;
typedef struct oxr_stream oxr_stream;
typedef struct oxr_buffer oxr_buffer;
typedef struct oxr_batch oxr_batch;
typedef struct oxr_join_table oxr_join_table;
typedef struct oxr_join_matches oxr_join_matches;
typedef struct oxr_agg_state oxr_agg_state;
typedef struct oxr_vector_index oxr_vector_index;

/* Arrow C Data Interface (https://arrow.apache.org/docs/format/CDataInterface.html) */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
  const char *format;
  const char *name;
  const char *metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema **children;
  struct ArrowSchema *dictionary;
  void (*release)(struct ArrowSchema *);
  void *private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void **buffers;
  struct ArrowArray **children;
  struct ArrowArray *dictionary;
  void (*release)(struct ArrowArray *);
  void *private_data;
};
#endif

/* ------------------------------------------------- device, memory, streams */

/* Whether device `ordinal` can run this backend (a gfx950 GPU with a HIP
 * runtime). On failure writes a reason into why[0..why_len). Mirrors
 * `CudaBackend::probe` (crates/oxidelake-device/src/cuda/mod.rs:293-335) and
 * the detector's typed "unavailable" reasons (detector.rs:86-127). */
int oxr_probe(int ordinal, char *why, size_t why_len);

/* Opens (once per process; later calls share it) device `ordinal`.
 * `CudaBackend::new` (cuda/mod.rs:344-365). */
int oxr_open(int ordinal, oxr_device **out);

/* `GpuBackend::memory_info` (backend.rs:52): total/free HBM bytes. */
int oxr_mem_info(oxr_device *dev, uint64_t *total_bytes, uint64_t *free_bytes,
                 int *pinned_supported);

/* Bytes the backend's HBM arena holds (in use + cached for reuse). */
int oxr_arena_info(oxr_device *dev, uint64_t *in_use, uint64_t *cached);

/* `GpuBackend::alloc_device` (backend.rs:54-55): zeroed, 256-byte aligned
 * (the trait asks for 128). */
int oxr_alloc_device(oxr_device *dev, size_t bytes, oxr_buffer **out);
/* `GpuBackend::alloc_pinned_host` (backend.rs:57-59): page-locked host
 * memory (`HostBuffer::class()` is always pinned here). */
int oxr_alloc_pinned(oxr_device *dev, size_t bytes, oxr_buffer **out);
void *oxr_buffer_ptr(const oxr_buffer *buf);
size_t oxr_buffer_len(const oxr_buffer *buf);
void oxr_buffer_free(oxr_buffer *buf);

/* `GpuBackend::create_stream` (backend.rs:62). Default priority. */
int oxr_stream_create(oxr_device *dev, oxr_stream **out);
/* Borrows an existing hipStream_t (e.g. a framework's current stream);
 * destroying the wrapper leaves the stream alone. NULL = the null stream. */
int oxr_stream_wrap(oxr_device *dev, void *hip_stream, oxr_stream **out);
void oxr_stream_destroy(oxr_stream *s);
/* `GpuBackend::synchronize` (backend.rs:81). */
int oxr_stream_sync(oxr_stream *s);
/* `GpuBackend::copy_h2d` / `copy_d2h` (backend.rs:65-79), async on `s`: the
 * host side must stay valid until the stream is synchronized. */
int oxr_copy_h2d(oxr_stream *s, const void *src, oxr_buffer *dst, size_t bytes);
int oxr_copy_d2h(oxr_stream *s, const oxr_buffer *src, void *dst, size_t bytes);

/* ------------------------------------------------------------------ batches */

/* `GpuBackend::upload` (backend.rs:84-88): a host RecordBatch exported as an
 * Arrow struct array. Columns of a type the device cannot hold, or a vector
 * column with a NULL element in a valid row (columns.rs:211-241), are
 * OXR_EUNSUPPORTED ("device.column_type" / "device.vector_null_element").
 * The caller keeps ownership of schema/array. The buffers travel through the
 * stream's double-buffered pinned staging ring (knobs transfer.staging,
 * transfer.chunk_kib): the reference copies from pageable memory
 * (cuda/mod.rs:445-498); the ring overlaps each chunk's host copy with the
 * other chunk's DMA. Blocks until the copies are done. */
int oxr_upload(oxr_stream *s, const struct ArrowSchema *schema,
               const struct ArrowArray *array, oxr_batch **out);

/* `GpuBackend::download` (backend.rs:91-95): exports the batch as an Arrow
 * struct array with host buffers; the caller releases both structs. Through
 * the same pinned ring as oxr_upload. Blocks until the copies are done. */
int oxr_download(oxr_stream *s, const oxr_batch *batch,
                 struct ArrowSchema *out_schema, struct ArrowArray *out_array);

/* A batch over device memory the caller owns (no copy). `names` may be NULL.
 * As for every batch, the memory must not change while the batch lives:
 * statistics measured on a column (an aggregation key's value range) are
 * kept with the batch. */
int oxr_batch_wrap(oxr_device *dev, const oxr_column *columns,
                   const char *const *names, int32_t ncols, uint64_t rows,
                   oxr_batch **out);
uint64_t oxr_batch_rows(const oxr_batch *b);
int32_t oxr_batch_ncols(const oxr_batch *b);
/* A view of column i (device pointers stay owned by the batch). */
int oxr_batch_column(const oxr_batch *b, int32_t i, oxr_column *out);
const char *oxr_batch_name(const oxr_batch *b, int32_t i);
void oxr_batch_free(oxr_batch *b);
/* Frees n batches (NULL entries skipped): one call for the many batches a
   multi-batch filter returns. No reference counterpart (the reference frees
   each batch on its own). */
void oxr_batch_free_many(oxr_batch *const *batches, uint64_t n);

/* ------------------------------------------------------------ expressions */
/*
 * Scalar expressions, postfix (children before parent). The plan vocabulary of
 * the reference (crates/oxidelake-core/src/params.rs:12-139) has
 * `Compare{column, op, literal}` and `And`; this grammar is a superset that
 * adds `Or`, Int32/Float32 literals, arithmetic (+ - * / %) and
 * expression-valued comparison operands, with DataFusion's numeric coercion:
 * Float64 if either side is Float64, else Float32 if either is Float32, else
 * Int64 if either is Int64, else Int32. Integer + - * wrap; integer / and %
 * truncate toward zero and fail (OXR_EEXECUTION) on a zero divisor.
 * Comparisons of floats use DataFusion 55's order: -0.0 read as +0.0, then
 * IEEE totalOrder (NaN above +inf, a sign-bit NaN below -inf).
 * A comparison with a NULL operand is NULL and a NULL predicate drops the row;
 * And/Or are Kleene (for trees without Not this is "NULL leaf -> false").
 */
enum {
  OXR_X_COLUMN = 1,  /* push column `column` */
  OXR_X_LITERAL = 2, /* push literal of `type` (lit_i for ints, lit_f for floats) */
  OXR_X_ADD = 3,
  OXR_X_SUB = 4,
  OXR_X_MUL = 5,
  OXR_X_DIV = 6,
  OXR_X_MOD = 7,
  OXR_X_CMP = 8,     /* pop b, a; push (a `op` b) as a boolean */
  OXR_X_AND = 9,
  OXR_X_OR = 10
};
enum { OXR_EQ = 1, OXR_LT = 2, OXR_LTEQ = 3, OXR_GT = 4, OXR_GTEQ = 5, OXR_NOTEQ = 6 };

typedef struct oxr_expr_node {
  int32_t kind;   /* OXR_X_* */
  int32_t type;   /* OXR_X_LITERAL: OXR_INT32/INT64/FLOAT32/FLOAT64 */
  int32_t column; /* OXR_X_COLUMN: input column index */
  int32_t op;     /* OXR_X_CMP: OXR_EQ .. OXR_NOTEQ */
  int64_t lit_i;
  double lit_f;
} oxr_expr_node;

typedef struct oxr_expr {
  const oxr_expr_node *nodes;
  int32_t count;
} oxr_expr;

/* Capability hooks asked before any upload: `GpuBackend::supports_predicate`
 * (backend.rs:182-185), `supports_hash_join` (:189-191) and
 * `supports_aggregate` (:195-197). Each returns 1 when the device runs the
 * operation (oxr_filter_project for that predicate; oxr_join_* and
 * oxr_aggregate / oxr_agg_*), 0 when every call would be OXR_EUNSUPPORTED.
 * Host-only: they touch no device. */
int oxr_supports_predicate(const oxr_expr *predicate);
int oxr_supports_hash_join(void);
int oxr_supports_aggregate(void);

/* ------------------------------------------------------- filter + project */

/* `GpuBackend::filter_project` (backend.rs:98-102, args.rs:10-17): rows of
 * `input` where `predicate` is true, as one output column per projection
 * expression, in input-row order. A projection that is a single column keeps
 * its type (vectors included) and validity; an arithmetic projection is
 * NULL where any operand is. `predicate` NULL or empty = keep every row.
 * `names` may be NULL (then "p0", "p1", ...). */
int oxr_filter_project(oxr_stream *s, const oxr_batch *input,
                       const oxr_expr *predicate, const oxr_expr *projection,
                       int32_t nproj, const char *const *names, oxr_batch **out);

/* The same over many batches in one launch: `input` read as consecutive
 * batches of `batch_rows` rows (a multiple of 32; the last may be short),
 * each filtered and compacted on its own, as the reference's operator does
 * batch by batch (crates/oxidelake-compute/src/exec/filter.rs:182-210), but
 * without a launch, a host round trip and an allocation per batch. Writes
 * ceil(rows / batch_rows) result batches (in input order) to outs[], which
 * has room for max_batches; *nbatches is that count. The result batches share
 * their device buffers; each may be freed independently. */
int oxr_filter_project_batches(oxr_stream *s, const oxr_batch *input, uint64_t batch_rows,
                               const oxr_expr *predicate, const oxr_expr *projection,
                               int32_t nproj, const char *const *names, oxr_batch **outs,
                               uint64_t max_batches, uint64_t *nbatches);

/* Bytes of device scratch every oxr_filter_project_noalloc call on a device
 * uses (the array is per device, held until the process ends). */
#define OXR_NOALLOC_SCRATCH_BYTES (2u << 20)

/* No counterpart in the reference: oxr_filter_project for callers that
 * cannot allocate (serving loops and harnesses that forbid device
 * allocation, code that must not call hipMalloc). The same rows, values and
 * NULLs as oxr_filter_project, written into the caller's `outputs` (nproj
 * columns: the type and dim each projection gives, as oxr_filter_project's
 * result reports them, `len` = capacity >= input rows, values 16-byte
 * aligned, a validity bitmap (4-byte aligned) when the projection can be
 * NULL; the first ceil(rows/32) validity words are cleared by the call).
 * *out_rows = the rows kept. The launch's scratch (look-back words, ticket,
 * error and count words) is a static device array of
 * OXR_NOALLOC_SCRATCH_BYTES: one 8-byte word per tile (512 rows on the
 * interpreter kernel's default shape, 4096 on a generated kernel of 16 rows
 * per lane) plus 128 bytes; an input needing more is refused (OXR_EPLAN)
 * before anything runs (about 134M rows at 512-row tiles). No device
 * memory is allocated, and no code is compiled or loaded: the query's
 * generated kernel runs when an earlier oxr_filter_project of the same query
 * shape loaded it, the interpreter kernel built into the library otherwise
 * (the same results). Calls on one device share its scratch array and take
 * turns (a lock held until the call's work has finished), so they never
 * overlap; the call returns when its results are in `outputs`. */
int oxr_filter_project_noalloc(oxr_stream *s, const oxr_batch *input, const oxr_expr *predicate,
                               const oxr_expr *projection, int32_t nproj, const oxr_column *outputs,
                               uint64_t *out_rows);

/* The kernel source oxr_filter_project generates for this query shape (an
 * EXPLAIN of the device plan; no counterpart in the reference, whose CUDA
 * kernels are fixed per operator, cuda/ops.rs:143-373). Only the type, dim
 * and validity (NULL or not) of each of the `ncols` columns are read; no
 * device is needed. With `compile` set it is also compiled for gfx950.
 * The text is the tile kernel's module followed, for shapes the staged
 * kernel takes (knob filter.staged), by the staged kernel's module.
 * Writes the NUL-terminated source to buf (truncated to cap-1 bytes) and its
 * full length to *len. OXR_EUNSUPPORTED: the query runs on the interpreter
 * kernel instead of a generated one; OXR_EPLAN: a knob asks for a staged
 * shape the kernel does not take. */
int oxr_filter_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_expr *predicate,
                             const oxr_expr *projection, int32_t nproj, int32_t compile,
                             char *buf, size_t cap, size_t *len);

/* -------------------------------------------------------------- hash join */

/* `GpuBackend::prepare_join_build` (backend.rs:129-142, args.rs:21-31): the
 * build side's index of every distinct non-NULL key's run of build rows, in
 * build order within a key, built on the device. `key` is an Int64 (or Int32)
 * column of `build`. */
int oxr_join_build(oxr_stream *s, const oxr_batch *build, int32_t key,
                   oxr_join_table **out);
/* Distinct keys and indexed build rows of a table. */
int oxr_join_table_info(const oxr_join_table *t, uint64_t *distinct_keys,
                        uint64_t *rows, int *dense);
void oxr_join_table_free(oxr_join_table *t);

/* `GpuBackend::probe_join` (backend.rs:153-163, args.rs:36-45): each probe
 * row's run in the table, numbered in probe-row order then build order
 * (backend.rs:148-152). `*total` is the number of matches. */
int oxr_join_probe(oxr_stream *s, const oxr_join_table *table,
                   const oxr_batch *probe, int32_t key, oxr_join_matches **out,
                   uint64_t *total);
void oxr_join_matches_free(oxr_join_matches *m);

/* `GpuBackend::hash_join` (backend.rs:114-118, args.rs:49-62): matches
 * [first, first+rows) as the probe batch's columns then the build batch's,
 * or the `projection` subset (indices into that concatenation) when
 * nproj >= 0. */
int oxr_join_gather(oxr_stream *s, const oxr_batch *probe,
                    const oxr_batch *build, const oxr_join_table *table,
                    const oxr_join_matches *matches, uint64_t first,
                    uint64_t rows, const int32_t *projection, int32_t nproj,
                    oxr_batch **out);

/* ---------------------------------------------------------- aggregation */
/* `AggregateFunction` (params.rs:117-127) plus COUNT(*). */
enum { OXR_SUM = 1, OXR_COUNT = 2, OXR_MIN = 3, OXR_MAX = 4, OXR_COUNT_STAR = 5 };

typedef struct oxr_agg {
  int32_t func;    /* OXR_SUM .. OXR_COUNT_STAR */
  int32_t pad;
  oxr_expr input;  /* ignored for OXR_COUNT_STAR */
  const char *name; /* output column name; NULL = "SUM(col)" style */
} oxr_agg;

/* `AggregateSpec` (params.rs:142-149) generalised: group key column (Int64
 * or Int32; -1 = no grouping, one global group) and aggregates over
 * expressions. `filter` (may be NULL) is applied first, fusing a
 * GpuFilterExec into the aggregate. `groups_hint` is the distinct keys the
 * state is expected to see (0 = unknown): the group table is sized for it at
 * create. It never limits the table: every update first grows the table so
 * that each key the batch can add (at most one per row, per build row of a
 * join, and per value of the key's measured range) keeps the load at most
 * one half. An update that fails after some of its rows were applied (a
 * division by zero, a device error) leaves the state unusable: every later
 * update or finish on it is OXR_EEXECUTION, never OXR_EUNSUPPORTED, so a
 * caller never re-runs on the CPU rows the device already counted. */
typedef struct oxr_agg_spec {
  int32_t group_by;
  int32_t naggs;
  const oxr_agg *aggs;
  const oxr_expr *filter;
  uint64_t groups_hint;
} oxr_agg_spec;

/* `GpuBackend::aggregate` (backend.rs:166-170): one batch's partial, laid
 * out [key, agg...] sorted by key with the NULL group last, as
 * `GroupedAggregator::merge` accepts (kernels.rs:774-790; cuda/ops.rs:1024).
 * SUM(Int32/Int64) is Int64 and wraps; SUM(Float32/Float64) is Float64 from
 * +0.0; MIN/MAX keep the input type (floats in totalOrder); COUNT is Int64. */
int oxr_aggregate(oxr_stream *s, const oxr_batch *input,
                  const oxr_agg_spec *spec, oxr_batch **out);

/* Streaming form: group state stays resident in HBM across batches (no
 * per-batch host merge). The spec's column indices refer to every update's
 * batch. */
int oxr_agg_create(oxr_stream *s, const oxr_batch *schema_like,
                   const oxr_agg_spec *spec, oxr_agg_state **out);
int oxr_agg_update(oxr_stream *s, oxr_agg_state *st, const oxr_batch *input);
/* Fused inner join + aggregation: rows of `probe` whose `probe_key` matches
 * `table` (built over `build`) aggregate into groups keyed by `build`'s
 * column `build_group` (the state's group_by is ignored). Aggregate inputs
 * are expressions over `probe`'s columns. Nothing is materialised. */
int oxr_agg_update_join(oxr_stream *s, oxr_agg_state *st,
                        const oxr_join_table *table, const oxr_batch *build,
                        int32_t build_group, const oxr_batch *probe,
                        int32_t probe_key);
int oxr_agg_finish(oxr_stream *s, oxr_agg_state *st, oxr_batch **out);
void oxr_agg_free(oxr_agg_state *st);

/* The aggregation kernel source generated for this spec over columns of the
 * given types and nullability (no device needed; no counterpart in the
 * reference, whose aggregation kernels are fixed, cuda/ops.rs:623-1161).
 * probe_key >= 0: the kernel of oxr_agg_update_join with that probe key
 * column (a unique-key join). OXR_EUNSUPPORTED: the interpreter kernel runs
 * this shape. Output as oxr_filter_kernel_source. */
int oxr_aggregate_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_agg_spec *spec,
                                int32_t probe_key, int32_t compile, char *buf, size_t cap,
                                size_t *len);

/* ------------------------------------------------ batched aggregations */
/* No counterpart in the reference, whose operator runs one aggregation per
 * plan over its batches (crates/oxidelake-compute/src/operator.rs:994-1038).
 * For a server answering many aggregation requests over one HBM-resident
 * table: requests of one prepared shape -- the same filter and aggregates
 * over one global group (group_by -1), differing only in literal values --
 * run as one pass over the table (each tile's rows loaded once, every
 * request's filter and aggregates evaluated on them), and the chain of a
 * run (the literals copied from pinned host memory, the scan, a final fold,
 * the results copied back to pinned memory) is captured once per batch size
 * into a hipGraph and replayed (knob agg.batch_graph; 0 launches it
 * directly). No device memory is allocated per run.
 * Every result is that of oxr_aggregate with the request's spec (one row:
 * COUNT 0 and the rest NULL over no selected row), FP64 sums compensated;
 * a request's additions run in a fixed order, so its result is the same
 * bits alone, in any batch, and from run to run. */
typedef struct oxr_agg_batch oxr_agg_batch;

/* Prepares `spec` over `table` (kept alive by the batch object; its values
 * must not change while it lives). *max_requests = the requests one run
 * takes (16 for up to 3 aggregates, fewer for more). OXR_EUNSUPPORTED: a
 * group key, or a shape outside the generated kernels. */
int oxr_agg_batch_prepare(oxr_stream *s, const oxr_batch *table, const oxr_agg_spec *spec,
                          int32_t *max_requests, oxr_agg_batch **out);
/* The output columns: their count, types (as oxr_aggregate's result
 * columns) and names (valid while the batch object lives). */
int oxr_agg_batch_columns(const oxr_agg_batch *b, int32_t *naggs, int32_t *types, const char **names);
/* Runs n (0..max_requests) requests: requests[r] is the prepared spec with
 * request r's literal values (OXR_EPLAN otherwise). Writes n * naggs values
 * (row-major; Int64 and Float64 as their 64 bits, Int32 and Float32 in the
 * low 32 bits), n * naggs valid flags (0 = NULL) and n statuses (OXR_OK, or
 * OXR_EEXECUTION for a request whose filter or inputs divided by zero in a
 * row: that request fails alone). Runs on one batch object take turns;
 * returns when the results are written. */
int oxr_agg_batch_run(oxr_stream *s, oxr_agg_batch *b, const oxr_agg_spec *requests, int32_t n,
                      int64_t *values, uint8_t *valid, int32_t *status);
void oxr_agg_batch_free(oxr_agg_batch *b);
/* The batched kernels' generated source for `spec` (as
 * oxr_aggregate_kernel_source; no device needed). */
int oxr_agg_batch_kernel_source(const oxr_column *cols, int32_t ncols, const oxr_agg_spec *spec,
                                int32_t compile, char *buf, size_t cap, size_t *len);

/* --------------------------------------------------------- vector distance */
enum { OXR_L2 = 1, OXR_COSINE = 2 }; /* DistanceMetric, params.rs:151-157 */

/* `GpuBackend::vector_distance` (backend.rs:173-177, args.rs:75-86): the
 * input columns plus an appended nullable Float32 distance column (validity
 * = the vector column's). L2 = sqrt(sum (a-b)^2); cosine = 1 - dot/(|a||b|),
 * NaN when a norm is zero (kernels.rs:908-927). */
int oxr_vector_distance(oxr_stream *s, const oxr_batch *input, int32_t column,
                        const float *query, int32_t dim, int32_t metric,
                        const char *output_name, oxr_batch **out);

/* Top-k (new: the reference sorts on the CPU, embedded.rs:24-25). For each
 * of `nq` queries (row-major nq x dim, host memory), the k nearest non-NULL
 * rows of the vector column as (distance ascending, row ascending on ties):
 * out_rows/out_dist are nq*k host arrays; missing entries are row -1. */
int oxr_vector_topk(oxr_stream *s, const oxr_batch *input, int32_t column,
                    const float *queries, int32_t nq, int32_t dim,
                    int32_t metric, int32_t k, int64_t *out_rows,
                    float *out_dist);

/* A search index over a vector column (new, like top-k): row norms and the
 * rows in the scan's precision, built once on the device and kept resident.
 * OXR_VEC_FP32 scans the FP32 rows themselves; OXR_VEC_BF16 a BF16 copy;
 * OXR_VEC_FP8 an e4m3 copy with one scale per row (max |x| / 448). The index
 * keeps the batch's buffers alive; the batch may be freed first. */
enum { OXR_VEC_FP32 = 1, OXR_VEC_BF16 = 2, OXR_VEC_FP8 = 3 };
int oxr_vector_index_create(oxr_stream *s, const oxr_batch *input, int32_t column,
                            int32_t metric, int32_t precision,
                            oxr_vector_index **out);
/* Top-k of `nq` queries (row-major nq x dim, host memory): the scan keeps
 * the `candidates` best rows per query (at least k, at most 1024) by the
 * index's precision, then ranks them by exact FP32 distance (ties: smaller
 * row first). out_rows/out_dist are nq*k host arrays; rows -1 where fewer
 * than k rows qualify. NULL rows and rows whose distance is NaN never
 * qualify. With OXR_VEC_FP32 and candidates = k this is the exact path. */
int oxr_vector_index_search(oxr_stream *s, const oxr_vector_index *index,
                            const float *queries, int32_t nq, int32_t k,
                            int32_t candidates, int64_t *out_rows,
                            float *out_dist);
void oxr_vector_index_free(oxr_vector_index *index);

/* --------------------------------------------------------- synthetic tables */
/* Seeded synthetic columns written straight into HBM, for benchmarks and
 * tests (the reference's counterpart is its `gen-data` table generator,
 * crates/oxidelake-storage/src/datagen.rs:89-182, which writes Parquet on
 * the host). Every value is a pure function of (seed, column index, row):
 * SplitMix64 draws, so a table is reproducible at any size. */
enum {
  OXR_GEN_UNIFORM_INT = 1,   /* integers in [a, b) */
  OXR_GEN_UNIFORM_REAL = 2,  /* reals in [a, b) */
  OXR_GEN_NORMAL = 3,        /* N(a, b^2) */
  OXR_GEN_LOGNORMAL = 4,     /* exp(N(a, b^2)) */
  OXR_GEN_ZIPF = 5,          /* Zipf exponent a over ranks 1..b, written as rank-1 */
  OXR_GEN_SEQUENCE = 6,      /* a + row */
  OXR_GEN_PERMUTATION = 7,   /* a + a random permutation of 0..rows-1 */
  OXR_GEN_MIXTURE = 8,       /* vectors: one of a N(0,1) centroids + N(0, b^2) noise */
  OXR_GEN_MIXTURE_NORMALIZED = 9 /* the same, L2-normalised */
};
typedef struct oxr_gen_spec {
  int32_t type;      /* OXR_INT32 .. OXR_F32_VECTOR */
  int32_t dist;      /* OXR_GEN_* */
  int32_t dim;       /* vector size for OXR_F32_VECTOR */
  int32_t pad;
  double a, b;       /* distribution parameters */
  double null_frac;  /* probability a row is NULL (0 = no validity bitmap) */
  uint64_t seed;
} oxr_gen_spec;
int oxr_generate(oxr_stream *s, const oxr_gen_spec *cols, const char *const *names,
                 int32_t ncols, uint64_t rows, oxr_batch **out);

/* ------------------------------------------------------------------ tuning */
/* Launch shapes and table sizes the autotuner picks (no counterpart in the
 * reference, whose CUDA launches are fixed at 256 threads, cuda/ops.rs:42).
 * Knobs: "filter.v" / "filter.g" (rows per lane, row groups per lane of the
 * filter kernel), "agg.grid" (blocks per CU of the aggregate kernel),
 * "vector.rows" (rows per block of the vector scan). A negative value resets
 * a knob to its default; oxr_tune_get reports -1 for a knob never set. */
int oxr_tune_set(const char *knob, int64_t value);
int oxr_tune_get(const char *knob, int64_t *value);

/* ---------------------------------------------------------------- profiling */
/* Per-kernel device time from HIP events recorded on the launching stream
 * around every kernel launch while enabled (off by default; benchmarks turn
 * it on). No counterpart in the reference: it stands in for its per-operator
 * elapsed_ns telemetry (crates/oxidelake-core/src/telemetry.rs:37-64) at
 * kernel granularity. */
void oxr_profile_enable(int on);
void oxr_profile_reset(void);
/* Total milliseconds and launch count of `kernel` since the last reset
 * (synchronizes on the recorded events). */
int oxr_profile_get(const char *kernel, double *total_ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* OXIDE_ROCM_H */

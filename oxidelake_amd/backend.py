"""`RocmBackend`: the object-safe hardware trait of the reference, over the C ABI.

Mirrors `trait GpuBackend` (crates/oxidelake-device/src/backend.rs:44-198)
method for method; every call goes to liboxide_rocm.so (include/oxide_rocm.h).
Errors mirror `EngineError` (crates/oxidelake-core/src/error.rs:17-73):
`Unsupported` is the per-batch fallback signal, every other error fails the
query. Host batches are pyarrow RecordBatches (the reference's arrow-rs
RecordBatch), moved through the Arrow C Data Interface.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import pyarrow as pa

from . import plan
from ._capi import (
    EALLOCATION,
    EDEVICE,
    EEXECUTION,
    EFORMAT,
    EPLAN,
    EUNSUPPORTED,
    Agg,
    AggSpec,
    ArrowArray,
    ArrowSchema,
    Column,
    GenSpec,
    lib,
)
from .plan import AggregateSpec, DistanceMetric, Encoded, ExprT


# ------------------------------------------------------------------ errors
class EngineError(Exception):
    kind = "Execution"

    def is_unsupported(self) -> bool:
        return False


class Unsupported(EngineError):
    kind = "Unsupported"

    def __init__(self, feature: str, detail: str):
        super().__init__(f"unsupported: {feature} ({detail})")
        self.feature = feature
        self.detail = detail

    def is_unsupported(self) -> bool:
        return True


class DeviceError(EngineError):
    kind = "Device"


class ExecutionError(EngineError):
    kind = "Execution"


class PlanError(EngineError):
    kind = "Plan"


class FormatError(EngineError):
    kind = "Format"


class AllocationError(EngineError):
    kind = "Allocation"


def check(rc: int) -> None:
    if rc == 0:
        return
    msg = (lib.oxr_last_error() or b"").decode()
    feat = (lib.oxr_last_feature() or b"").decode()
    if rc == EUNSUPPORTED:
        raise Unsupported(feat or "rocm", msg)
    cls = {
        EDEVICE: DeviceError,
        EEXECUTION: ExecutionError,
        EPLAN: PlanError,
        EFORMAT: FormatError,
        EALLOCATION: AllocationError,
    }.get(rc, EngineError)
    raise cls(f"rocm device error: {msg}" if rc == EDEVICE else msg)


# ------------------------------------------------------------------ handles
class Stream:
    """`DeviceStream` (handles.rs): an owned HIP stream or a borrowed one."""

    def __init__(self, handle: int, backend: "RocmBackend"):
        self.handle = handle
        self.backend = backend

    def synchronize(self) -> None:
        check(lib.oxr_stream_sync(self.handle))

    def __del__(self):
        if getattr(self, "handle", None):
            lib.oxr_stream_destroy(self.handle)
            self.handle = None


_TYPESTR = {plan.INT32: "<i4", plan.INT64: "<i8", plan.FLOAT32: "<f4", plan.FLOAT64: "<f8",
            plan.F32_VECTOR: "<f4"}
_WIDTH = {plan.INT32: 4, plan.INT64: 8, plan.FLOAT32: 4, plan.FLOAT64: 8}


class DeviceBatch:
    """`DeviceBatch` (handles.rs:130): a batch resident in HBM. `owned`
    False: a view whose handle another object frees (DeviceBatches)."""

    def __init__(self, handle: int, keepalive=None, owned: bool = True):
        self.handle = handle
        self._keep = keepalive
        self._owned = owned

    @property
    def num_rows(self) -> int:
        return int(lib.oxr_batch_rows(self.handle))

    @property
    def num_columns(self) -> int:
        return int(lib.oxr_batch_ncols(self.handle))

    def name(self, i: int) -> str:
        return lib.oxr_batch_name(self.handle, i).decode()

    def column(self, i: int) -> Column:
        c = Column()
        check(lib.oxr_batch_column(self.handle, i, C.byref(c)))
        return c

    def __del__(self):
        if getattr(self, "handle", None) and getattr(self, "_owned", True):
            lib.oxr_batch_free(self.handle)
        self.handle = None


class DeviceBatches(Sequence):
    """The batches of one multi-batch call, freed together
    (oxr_batch_free_many) when the last reference goes; indexing gives
    views that keep the set alive."""

    def __init__(self, handles, n: int):
        self._h = handles  # ctypes array of oxr_batch*
        self._n = n

    def __len__(self) -> int:
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self._n))]
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError(i)
        return DeviceBatch(self._h[i], keepalive=self, owned=False)

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib.oxr_batch_free_many(self._h, self._n)
            self._h = None


class JoinTable:
    """`DeviceJoinTable` (handles.rs:224)."""

    def __init__(self, handle: int):
        self.handle = handle

    def info(self) -> tuple[int, int, bool]:
        d, r, dense = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib.oxr_join_table_info(self.handle, C.byref(d), C.byref(r), C.byref(dense)))
        return int(d.value), int(r.value), bool(dense.value)

    def __del__(self):
        if getattr(self, "handle", None):
            lib.oxr_join_table_free(self.handle)
            self.handle = None


class JoinMatches:
    """`DeviceJoinMatches` (handles.rs:303): one probe batch's matches."""

    def __init__(self, handle: int, total: int, probe_rows: int):
        self.handle = handle
        self.total = total
        self.probe_rows = probe_rows

    def __len__(self) -> int:
        return self.total

    def __del__(self):
        if getattr(self, "handle", None):
            lib.oxr_join_matches_free(self.handle)
            self.handle = None


@dataclass
class MemoryInfo:
    total_bytes: int
    free_bytes: int
    pinned_supported: bool

    @property
    def used_bytes(self) -> int:
        return max(0, self.total_bytes - self.free_bytes)


class DeviceBuffer:
    """`DeviceBuffer` (crates/oxidelake-device/src/handles.rs): zeroed HBM
    from the backend's arena, freed when dropped."""

    def __init__(self, handle: int):
        self.handle = handle

    def __len__(self) -> int:
        return int(lib.oxr_buffer_len(self.handle))

    @property
    def ptr(self) -> int:
        return int(lib.oxr_buffer_ptr(self.handle) or 0)

    def free(self) -> None:
        if self.handle:
            lib.oxr_buffer_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer(DeviceBuffer):
    """`HostBuffer` (crates/oxidelake-memory/src/pinned.rs:24-35): page-locked
    host memory; `class_` is "pinned" (the class the trait's `class()` reports)."""

    class_ = "pinned"

    def array(self) -> np.ndarray:
        """The buffer's bytes as a writable numpy view."""
        n = len(self)
        return np.ctypeslib.as_array((C.c_uint8 * n).from_address(self.ptr)) if n else np.empty(0, np.uint8)


def filter_kernel_source(columns: Sequence[tuple], predicate: ExprT | None, projection: Sequence,
                         compile: bool = False) -> str:
    """The kernel source the filter generates for this query shape over
    columns given as (type, nullable) or (type, dim, nullable); with
    `compile`, it is also compiled for gfx950. No device is needed.
    Raises Unsupported when the interpreter kernel would run the query."""
    cols = (Column * max(1, len(columns)))()
    for i, c in enumerate(columns):
        typ, dim, nullable = (c[0], 0, c[1]) if len(c) == 2 else c
        cols[i].type = typ
        cols[i].dim = dim
        cols[i].validity = 1 if nullable else None  # only NULL or not is read
    exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
    pred = Encoded(predicate)
    parr, penc = plan.encode_many(exprs)
    n = C.c_size_t()
    check(lib.oxr_filter_kernel_source(cols, len(columns), pred.ref(), parr, len(exprs), 0, None, 0,
                                       C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib.oxr_filter_kernel_source(cols, len(columns), pred.ref(), parr, len(exprs),
                                       1 if compile else 0, buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def aggregate_kernel_source(columns: Sequence[tuple], spec: AggregateSpec, probe_key: int = -1,
                            compile: bool = False) -> str:
    """The aggregation kernel generated for `spec` over columns given as
    (type, nullable) (probe_key >= 0: the fused join's kernel). No device is
    needed. Raises Unsupported when the interpreter kernel runs the shape."""
    cols = (Column * max(1, len(columns)))()
    for i, c in enumerate(columns):
        typ, dim, nullable = (c[0], 0, c[1]) if len(c) == 2 else c
        cols[i].type = typ
        cols[i].dim = dim
        cols[i].validity = 1 if nullable else None
    cspec, keep = _agg_spec(spec)
    n = C.c_size_t()
    check(lib.oxr_aggregate_kernel_source(cols, len(columns), C.byref(cspec), probe_key, 0, None, 0,
                                          C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib.oxr_aggregate_kernel_source(cols, len(columns), C.byref(cspec), probe_key,
                                          1 if compile else 0, buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def _cstrs(names: Optional[Sequence[str]], n: int):
    if names is None:
        return None, None
    raw = [s.encode() for s in names]
    arr = (C.c_char_p * max(1, n))(*raw)
    return arr, raw


class RocmBackend:
    """The MI355X backend behind the reference's `GpuBackend` trait."""

    def __init__(self, ordinal: int = 0):
        why = C.create_string_buffer(256)
        rc = lib.oxr_probe(ordinal, why, 256)
        if rc != 0:
            raise Unsupported("rocm.probe", why.value.decode() or "no gfx950 device")
        dev = C.c_void_p()
        check(lib.oxr_open(ordinal, C.byref(dev)))
        self.dev = dev.value
        self.ordinal = ordinal

    # -- identity (backend.rs:46-53)
    def kind(self) -> str:
        return "rocm"

    def device_id(self) -> int:
        return self.ordinal

    def memory_info(self) -> MemoryInfo:
        t, f, p = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib.oxr_mem_info(self.dev, C.byref(t), C.byref(f), C.byref(p)))
        return MemoryInfo(int(t.value), int(f.value), bool(p.value))

    # -- memory (backend.rs:54-79)
    def alloc_device(self, nbytes: int) -> DeviceBuffer:
        """backend.rs:54-55: zeroed device memory."""
        out = C.c_void_p()
        check(lib.oxr_alloc_device(self.dev, int(nbytes), C.byref(out)))
        return DeviceBuffer(out.value)

    def alloc_pinned_host(self, nbytes: int) -> HostBuffer:
        """backend.rs:57-59: page-locked, zeroed host memory."""
        out = C.c_void_p()
        check(lib.oxr_alloc_pinned(self.dev, int(nbytes), C.byref(out)))
        return HostBuffer(out.value)

    def copy_h2d(self, stream: Stream, src: HostBuffer, dst: DeviceBuffer) -> None:
        """backend.rs:65-71: all of `src` into the start of `dst`, async on
        `stream` (a copy larger than `dst` is an ExecutionError)."""
        check(lib.oxr_copy_h2d(stream.handle, src.ptr, dst.handle, len(src)))

    def copy_d2h(self, stream: Stream, src: DeviceBuffer, dst: HostBuffer) -> None:
        """backend.rs:73-79: the start of `src` into all of `dst`, async on
        `stream`."""
        check(lib.oxr_copy_d2h(stream.handle, src.handle, dst.ptr, len(dst)))

    # -- streams (backend.rs:62, 81)
    def create_stream(self) -> Stream:
        s = C.c_void_p()
        check(lib.oxr_stream_create(self.dev, C.byref(s)))
        return Stream(s.value, self)

    def wrap_stream(self, hip_stream: int) -> Stream:
        s = C.c_void_p()
        check(lib.oxr_stream_wrap(self.dev, C.c_void_p(hip_stream), C.byref(s)))
        return Stream(s.value, self)

    def synchronize(self, stream: Stream) -> None:
        stream.synchronize()

    # -- batches (backend.rs:86-97)
    def upload(self, stream: Stream, batch: pa.RecordBatch) -> DeviceBatch:
        schema, array = ArrowSchema(), ArrowArray()
        batch._export_to_c(C.addressof(array), C.addressof(schema))
        out = C.c_void_p()
        try:
            check(lib.oxr_upload(stream.handle, C.byref(schema), C.byref(array), C.byref(out)))
        finally:
            _release(array)
            _release(schema)
        return DeviceBatch(out.value)

    def download(self, stream: Stream, batch: DeviceBatch) -> pa.RecordBatch:
        schema, array = ArrowSchema(), ArrowArray()
        check(lib.oxr_download(stream.handle, batch.handle, C.byref(schema), C.byref(array)))
        return pa.RecordBatch._import_from_c(C.addressof(array), C.addressof(schema))

    def wrap(self, columns: Sequence[tuple], names: Optional[Sequence[str]] = None,
             rows: Optional[int] = None) -> DeviceBatch:
        """A batch over device memory the caller owns (no copy): each column is
        (values, validity_or_None, oxr_type[, dim]) where values/validity are
        objects with `data_ptr()` and `shape` (e.g. framework tensors)."""
        cols = (Column * max(1, len(columns)))()
        keep = []
        nrows = rows
        for i, spec in enumerate(columns):
            t, valid, typ = spec[0], spec[1], spec[2]
            dim = spec[3] if len(spec) > 3 else (t.shape[1] if typ == plan.F32_VECTOR else 0)
            n = t.shape[0]
            nrows = n if nrows is None else nrows
            cols[i].type = typ
            cols[i].dim = dim
            cols[i].len = n
            cols[i].values = t.data_ptr()
            cols[i].validity = valid.data_ptr() if valid is not None else None
            keep.extend([t, valid])
        arr, raw = _cstrs(names, len(columns))
        out = C.c_void_p()
        check(lib.oxr_batch_wrap(self.dev, cols, arr, len(columns), nrows or 0, C.byref(out)))
        return DeviceBatch(out.value, keepalive=(keep, raw))

    def generate(self, stream: Stream, columns: Sequence, rows: int) -> DeviceBatch:
        """A seeded synthetic table written straight into HBM (oxr_generate).
        `columns` are datagen.GenColumn specs."""
        specs = (GenSpec * max(1, len(columns)))()
        for i, c in enumerate(columns):
            specs[i] = GenSpec(c.type, c.dist, c.dim, 0, float(c.a), float(c.b), float(c.null_frac),
                               int(c.seed) & 0xFFFFFFFFFFFFFFFF)
        arr, raw = _cstrs([c.name for c in columns], len(columns))
        out = C.c_void_p()
        check(lib.oxr_generate(stream.handle, specs, arr, len(columns), rows, C.byref(out)))
        return DeviceBatch(out.value)

    # -- operators
    def supports_predicate(self, predicate: ExprT | None) -> bool:
        """backend.rs:182-185."""
        enc = Encoded(predicate)
        return bool(lib.oxr_supports_predicate(enc.ref()))

    def supports_hash_join(self) -> bool:
        """backend.rs:189-191."""
        return bool(lib.oxr_supports_hash_join())

    def supports_aggregate(self) -> bool:
        """backend.rs:195-197."""
        return bool(lib.oxr_supports_aggregate())

    def filter_project(self, stream: Stream, batch: DeviceBatch, predicate: ExprT | None,
                       projection: Sequence, names: Optional[Sequence[str]] = None) -> DeviceBatch:
        """backend.rs:98-102. `projection` items are column indices (the
        reference's form) or expressions."""
        exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        if names is None:
            names = [batch.name(p.index) if isinstance(p, plan.Col) else f"p{i}"
                     for i, p in enumerate(exprs)]
        pred = Encoded(predicate)
        parr, penc = plan.encode_many(exprs)
        narr, raw = _cstrs(names, len(exprs))
        out = C.c_void_p()
        check(lib.oxr_filter_project(stream.handle, batch.handle, pred.ref(), parr, len(exprs),
                                     narr, C.byref(out)))
        return DeviceBatch(out.value)

    def prepare_filter(self, predicate: ExprT | None, projection: Sequence,
                       names: Optional[Sequence[str]] = None) -> "PreparedFilter":
        """A filter + projection encoded once, to run over many batches."""
        return PreparedFilter(predicate, projection, names)

    def filter_project_batches(self, stream: Stream, batch: DeviceBatch, batch_rows: int,
                               predicate: ExprT | None, projection: Sequence,
                               names: Optional[Sequence[str]] = None) -> "DeviceBatches":
        """`batch` read as consecutive batches of `batch_rows` rows, each
        filtered on its own, all in one launch (oxr_filter_project_batches)."""
        exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        if names is None:
            names = [batch.name(p.index) if isinstance(p, plan.Col) else f"p{i}"
                     for i, p in enumerate(exprs)]
        pred = Encoded(predicate)
        parr, penc = plan.encode_many(exprs)
        narr, raw = _cstrs(names, len(exprs))
        cap = max(1, -(-batch.num_rows // batch_rows))
        outs = (C.c_void_p * cap)()
        nb = C.c_uint64()
        check(lib.oxr_filter_project_batches(stream.handle, batch.handle, batch_rows, pred.ref(), parr,
                                             len(exprs), narr, outs, cap, C.byref(nb)))
        return DeviceBatches(outs, nb.value)

    def filter_project_noalloc(self, stream: Stream, batch: DeviceBatch, predicate: ExprT | None,
                               projection: Sequence, outputs: Sequence[tuple]) -> int:
        """oxr_filter_project_noalloc: the filter's rows written into caller
        device memory, `outputs` one (values_ptr, validity_ptr or 0, oxr_type,
        dim, capacity_rows) per projection; no device allocation inside.
        Returns the rows kept."""
        exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        if len(outputs) != len(exprs):
            raise ValueError(f"{len(exprs)} projections, {len(outputs)} output columns")
        cols = (Column * max(1, len(outputs)))()
        for i, (values, validity, typ, dim, cap) in enumerate(outputs):
            cols[i].type, cols[i].dim, cols[i].len = typ, dim, cap
            cols[i].values, cols[i].validity = values or None, validity or None
        pred = Encoded(predicate)
        parr, penc = plan.encode_many(exprs)
        rows = C.c_uint64()
        check(lib.oxr_filter_project_noalloc(stream.handle, batch.handle, pred.ref(), parr, len(exprs), cols,
                                             C.byref(rows)))
        return int(rows.value)

    def prepare_join_build(self, stream: Stream, build: DeviceBatch, key: int) -> JoinTable:
        """backend.rs:129-142 (the table is built on the device)."""
        out = C.c_void_p()
        check(lib.oxr_join_build(stream.handle, build.handle, key, C.byref(out)))
        return JoinTable(out.value)

    def probe_join(self, stream: Stream, probe: DeviceBatch, key: int, table: JoinTable) -> JoinMatches:
        """backend.rs:153-163."""
        out, total = C.c_void_p(), C.c_uint64()
        check(lib.oxr_join_probe(stream.handle, table.handle, probe.handle, key, C.byref(out),
                                 C.byref(total)))
        return JoinMatches(out.value, int(total.value), probe.num_rows)

    def hash_join(self, stream: Stream, left: DeviceBatch, right: DeviceBatch, table: JoinTable,
                  matches: JoinMatches, first: int, rows: int,
                  projection: Optional[Sequence[int]] = None) -> DeviceBatch:
        """backend.rs:114-118."""
        out = C.c_void_p()
        if projection is None:
            proj, n = None, -1
        else:
            proj = (C.c_int32 * max(1, len(projection)))(*projection)
            n = len(projection)
        check(lib.oxr_join_gather(stream.handle, left.handle, right.handle, table.handle,
                                  matches.handle, first, rows, proj, n, C.byref(out)))
        return DeviceBatch(out.value)

    def aggregate(self, stream: Stream, batch: DeviceBatch, spec: AggregateSpec) -> DeviceBatch:
        """backend.rs:166-170."""
        cspec, keep = _agg_spec(spec)
        out = C.c_void_p()
        check(lib.oxr_aggregate(stream.handle, batch.handle, C.byref(cspec), C.byref(out)))
        return DeviceBatch(out.value)

    def vector_distance(self, stream: Stream, batch: DeviceBatch, column: int,
                        query: Sequence[float], metric: DistanceMetric,
                        output_name: str) -> DeviceBatch:
        """backend.rs:173-177."""
        q = np.ascontiguousarray(np.asarray(query, dtype=np.float32))
        out = C.c_void_p()
        check(lib.oxr_vector_distance(stream.handle, batch.handle, column,
                                      q.ctypes.data_as(C.POINTER(C.c_float)), len(q), int(metric),
                                      output_name.encode(), C.byref(out)))
        return DeviceBatch(out.value)

    def prepare_agg_batch(self, stream: Stream, table: DeviceBatch, spec: AggregateSpec) -> "AggBatch":
        """Batched scalar aggregations of `spec`'s shape over `table`
        (oxr_agg_batch_prepare); see AggBatch."""
        return AggBatch(stream, table, spec)

    def aggregate_state(self, stream: Stream, like: DeviceBatch, spec: AggregateSpec) -> "AggState":
        """Groups resident in HBM across batches (oxr_agg_create): each
        `update` folds one batch of `like`'s schema, `finish` returns the
        partial laid out as `aggregate`'s."""
        return AggState(self, stream, like, spec)

    def vector_topk(self, stream: Stream, batch: DeviceBatch, column: int, queries: np.ndarray,
                    metric: DistanceMetric, k: int) -> tuple[np.ndarray, np.ndarray]:
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
        if q.ndim == 1:
            q = q[None, :]
        nq, dim = q.shape
        rows = np.empty((nq, k), dtype=np.int64)
        dist = np.empty((nq, k), dtype=np.float32)
        check(lib.oxr_vector_topk(stream.handle, batch.handle, column,
                                  q.ctypes.data_as(C.POINTER(C.c_float)), nq, dim, int(metric), k,
                                  rows.ctypes.data_as(C.POINTER(C.c_int64)),
                                  dist.ctypes.data_as(C.POINTER(C.c_float))))
        return rows, dist


class PreparedFilter:
    """The plan of a filter + projection encoded once (RocmBackend.
    prepare_filter): run and run_batches skip the per-call encoding."""

    def __init__(self, predicate: ExprT | None, projection: Sequence, names: Optional[Sequence[str]]):
        self.exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        self.names = names
        self._pred = Encoded(predicate)
        self._parr, self._penc = plan.encode_many(self.exprs)

    def _cnames(self, batch: DeviceBatch):
        names = self.names
        if names is None:
            names = [batch.name(p.index) if isinstance(p, plan.Col) else f"p{i}" for i, p in enumerate(self.exprs)]
        return _cstrs(names, len(self.exprs))

    def run(self, stream: Stream, batch: DeviceBatch) -> DeviceBatch:
        narr, raw = self._cnames(batch)
        out = C.c_void_p()
        check(lib.oxr_filter_project(stream.handle, batch.handle, self._pred.ref(), self._parr, len(self.exprs),
                                     narr, C.byref(out)))
        return DeviceBatch(out.value)

    def run_batches(self, stream: Stream, batch: DeviceBatch, batch_rows: int) -> "DeviceBatches":
        narr, raw = self._cnames(batch)
        cap = max(1, -(-batch.num_rows // batch_rows))
        outs = (C.c_void_p * cap)()
        nb = C.c_uint64()
        check(lib.oxr_filter_project_batches(stream.handle, batch.handle, batch_rows, self._pred.ref(), self._parr,
                                             len(self.exprs), narr, outs, cap, C.byref(nb)))
        return DeviceBatches(outs, nb.value)


class VectorIndex:
    """A resident search index over a vector column (oxr_vector_index_*):
    precision VEC_FP32 (exact), VEC_BF16 or VEC_FP8 (reduced-precision scan
    with an FP32 rerank of `candidates` rows per query)."""

    def __init__(self, stream: Stream, batch: DeviceBatch, column: int, metric: DistanceMetric,
                 precision: int):
        out = C.c_void_p()
        check(lib.oxr_vector_index_create(stream.handle, batch.handle, column, int(metric),
                                          int(precision), C.byref(out)))
        self.handle = out.value
        self.precision = precision

    def search(self, stream: Stream, queries: np.ndarray, k: int,
               candidates: int = 0) -> tuple[np.ndarray, np.ndarray]:
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        rows = np.empty((nq, k), dtype=np.int64)
        dist = np.empty((nq, k), dtype=np.float32)
        check(lib.oxr_vector_index_search(stream.handle, self.handle,
                                          q.ctypes.data_as(C.POINTER(C.c_float)), nq, k,
                                          max(candidates, k),
                                          rows.ctypes.data_as(C.POINTER(C.c_int64)),
                                          dist.ctypes.data_as(C.POINTER(C.c_float))))
        return rows, dist

    def free(self) -> None:
        if self.handle:
            lib.oxr_vector_index_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


VEC_FP32, VEC_BF16, VEC_FP8 = 1, 2, 3


class AggState:
    """Streaming aggregation state resident in HBM (oxr_agg_*)."""

    def __init__(self, backend: RocmBackend, stream: Stream, like: DeviceBatch, spec: AggregateSpec):
        self.cspec, self._keep = _agg_spec(spec)
        out = C.c_void_p()
        check(lib.oxr_agg_create(stream.handle, like.handle, C.byref(self.cspec), C.byref(out)))
        self.handle = out.value

    def update(self, stream: Stream, batch: DeviceBatch) -> None:
        check(lib.oxr_agg_update(stream.handle, self.handle, batch.handle))

    def update_join(self, stream: Stream, table: JoinTable, build: DeviceBatch, build_group: int,
                    probe: DeviceBatch, probe_key: int) -> None:
        check(lib.oxr_agg_update_join(stream.handle, self.handle, table.handle, build.handle,
                                      build_group, probe.handle, probe_key))

    def finish(self, stream: Stream) -> DeviceBatch:
        out = C.c_void_p()
        check(lib.oxr_agg_finish(stream.handle, self.handle, C.byref(out)))
        return DeviceBatch(out.value)

    def __del__(self):
        if getattr(self, "handle", None):
            lib.oxr_agg_free(self.handle)
            self.handle = None


class AggBatch:
    """Batched scalar aggregations over one resident table (oxr_agg_batch_*):
    requests of one prepared shape (the same filter and aggregates, no group
    key, only literal values differing) answered by one pass over the table,
    the run's chain replayed from a captured hipGraph. No counterpart in the
    reference (its operator runs one aggregation per plan,
    crates/oxidelake-compute/src/operator.rs:994-1038)."""

    _ARROW = {plan.INT32: pa.int32(), plan.INT64: pa.int64(), plan.FLOAT32: pa.float32(),
              plan.FLOAT64: pa.float64()}

    def __init__(self, stream: Stream, table: DeviceBatch, spec: AggregateSpec):
        cspec, keep = _agg_spec(spec)
        out, kmax = C.c_void_p(), C.c_int32()
        check(lib.oxr_agg_batch_prepare(stream.handle, table.handle, C.byref(cspec), C.byref(kmax),
                                        C.byref(out)))
        self.handle = out.value
        self.table = table  # the object keeps its buffers alive; this keeps the Python side too
        self.max_requests = int(kmax.value)
        na = C.c_int32()
        check(lib.oxr_agg_batch_columns(self.handle, C.byref(na), None, None))
        types = (C.c_int32 * max(1, na.value))()
        names = (C.c_char_p * max(1, na.value))()
        check(lib.oxr_agg_batch_columns(self.handle, C.byref(na), types, names))
        self.schema = pa.schema([pa.field(names[i].decode(), self._ARROW[types[i]]) for i in range(na.value)])

    def run(self, stream: Stream, specs: Sequence[AggregateSpec]) -> list:
        """One result per request: a one-row RecordBatch (as oxr_aggregate's
        global aggregate), or the ExecutionError of a request that divided by
        zero (it fails alone)."""
        n = len(specs)
        if n == 0:
            return []
        arr = (AggSpec * n)()
        keep = []
        for i, sp in enumerate(specs):
            cs, k = _agg_spec(sp)
            arr[i] = cs
            keep.append(k)
        na = len(self.schema)
        values = np.zeros(n * na, np.int64)
        valid = np.zeros(n * na, np.uint8)
        status = np.zeros(n, np.int32)
        check(lib.oxr_agg_batch_run(stream.handle, self.handle, arr, n,
                                    values.ctypes.data_as(C.POINTER(C.c_int64)),
                                    valid.ctypes.data_as(C.POINTER(C.c_uint8)),
                                    status.ctypes.data_as(C.POINTER(C.c_int32))))
        out = []
        for r in range(n):
            if status[r] != 0:
                out.append(ExecutionError(f"request {r}: division by zero"))
                continue
            cols = []
            for a, f in enumerate(self.schema):
                # the value's bits: 32-bit types in the low word (little endian: element 0)
                word = values[r * na + a:r * na + a + 1]
                v = word.view(np.dtype(f.type.to_pandas_dtype()))[:1]
                cols.append(pa.array(v, f.type, mask=np.array([valid[r * na + a] == 0])))
            out.append(pa.RecordBatch.from_arrays(cols, schema=self.schema))
        return out

    def free(self) -> None:
        if getattr(self, "handle", None):
            lib.oxr_agg_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        self.free()


def agg_batch_kernel_source(columns: Sequence[tuple], spec: AggregateSpec, compile: bool = False) -> str:
    """The batched-request kernels generated for `spec` over columns given as
    (type, nullable). No device is needed."""
    cols = (Column * max(1, len(columns)))()
    for i, c in enumerate(columns):
        typ, dim, nullable = (c[0], 0, c[1]) if len(c) == 2 else c
        cols[i].type = typ
        cols[i].dim = dim
        cols[i].validity = 1 if nullable else None
    cspec, keep = _agg_spec(spec)
    n = C.c_size_t()
    check(lib.oxr_agg_batch_kernel_source(cols, len(columns), C.byref(cspec), 0, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(lib.oxr_agg_batch_kernel_source(cols, len(columns), C.byref(cspec), 1 if compile else 0, buf,
                                          n.value + 1, C.byref(n)))
    return buf.value.decode()


def _agg_spec(spec: AggregateSpec):
    aggs = (Agg * max(1, len(spec.aggregates)))()
    keep = []
    for i, a in enumerate(spec.aggregates):
        enc = Encoded(a.input)
        keep.append(enc)
        aggs[i].func = int(a.func)
        aggs[i].input = enc.expr
        if a.name is not None:
            nm = a.name.encode()
            keep.append(nm)
            aggs[i].name = nm
        else:
            aggs[i].name = None
    flt = None
    if spec.filter is not None:
        fenc = Encoded(spec.filter)
        keep.append(fenc)
        flt = C.pointer(fenc.expr)
    cspec = AggSpec(spec.group_by, len(spec.aggregates), aggs, flt, spec.groups_hint)
    keep.append(aggs)
    return cspec, keep


def _release(struct) -> None:
    if struct.release:
        fn = C.CFUNCTYPE(None, C.c_void_p)(struct.release)
        fn(C.addressof(struct))

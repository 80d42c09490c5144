"""The caller side of the backend over host Arrow batches: the counterpart of
the reference's `GpuOperator` (crates/oxidelake-compute/src/operator.rs:
1-1145), which every `Gpu*Exec` drives one batch at a time.

- Only the columns a kernel reads travel (`SidePlan`, operator.rs:87-124): a
  filter's predicate and projected columns, a join's keys and projected
  columns, an aggregate's key and inputs, a distance's vector column.
- A projected column the device does not take (Utf8, Int8, a vector with a
  NULL element in a valid row) stays on the host; a row-id column travels
  instead and the host gathers the column at the ids the device kept
  (operator.rs:37-60, 150-154).
- The backend is asked before anything is uploaded (`supports_*`).
- A join's build side is uploaded and indexed once per operator
  (operator.rs:617-688) and every probe batch's output leaves in pieces of
  at most `batch_size` rows (operator.rs:705-806).
- A batch the backend declines (`Unsupported`) runs the caller's CPU path
  and is counted; a join piece declined mid-batch continues on the CPU path
  from the same match, since both number a probe's matches alike
  (operator.rs:758-775).

Two choices differ from the reference's operator on purpose:
- An aggregation's groups stay resident in one HBM table across batches
  (`AggState`) and come back once, at `finish`; the reference downloads a
  partial per batch and merges it on the host (operator.rs:994-1038). A
  declined batch is folded by the CPU path's own aggregator, which merges
  the device's groups at `finish`. `rows_out` is counted there, once.
- The CPU path is the caller's (`cpu`, see `CpuPath`): this package holds no
  CPU arithmetic. Without one, a declined batch raises its `Unsupported`.
"""
from __future__ import annotations

import logging
import threading
import time
from dataclasses import dataclass, field
from typing import Optional, Protocol, Sequence

import numpy as np
import pyarrow as pa

from . import plan
from .backend import DeviceBatch, ExecutionError, PlanError, RocmBackend, Unsupported
from .loader import device_type
from .plan import Aggregate, AggregateSpec, DistanceMetric

log = logging.getLogger(__name__)

ROW_ID = "__oxide_row"  # operator.rs:41-43


class CpuPath(Protocol):
    """What a declined batch runs instead (the reference's CPU kernels,
    crates/oxidelake-device/src/cpu/kernels.rs). Supplied by the caller."""

    def filter_project(self, batch: pa.RecordBatch, predicate, projection: Sequence,
                       names: Optional[Sequence[str]]) -> pa.RecordBatch: ...

    def join_matches(self, left: pa.RecordBatch, left_key: int, build: "JoinBuild"): ...

    def join_gather(self, left: pa.RecordBatch, build: "JoinBuild", matches, first: int, rows: int,
                    projection: Optional[Sequence[int]]) -> pa.RecordBatch: ...

    def aggregator(self, schema_batch: pa.RecordBatch, spec: AggregateSpec): ...

    def vector_distance(self, batch: pa.RecordBatch, column: int, query, metric: DistanceMetric,
                        output_name: str) -> pa.RecordBatch: ...


@dataclass
class OperatorStats:
    """The counters of one operator (oxidelake-core telemetry
    `OperatorStats`): batches, rows in and out, bytes each way, the batches
    that ran the CPU path, and wall time."""

    batches: int = 0
    rows_in: int = 0
    rows_out: int = 0
    bytes_h2d: int = 0
    bytes_d2h: int = 0
    fallback_batches: int = 0
    elapsed_s: float = 0.0
    fallbacks: dict = field(default_factory=dict)  # feature -> batches


# ------------------------------------------------------------ column plumbing
def _null_element_in_valid_row(arr: pa.Array) -> bool:
    """columns.rs:104-160: a vector row that is valid but holds a NULL
    element (the device format has one validity bit per row)."""
    if not pa.types.is_fixed_size_list(arr.type) or len(arr) == 0:
        return False
    size = arr.type.list_size
    child = arr.values.slice(arr.offset * size, len(arr) * size)
    if child.null_count == 0:
        return False
    bad = np.asarray(child.is_null()).reshape(len(arr), size).any(axis=1)
    return bool((bad & np.asarray(arr.is_valid())).any())


def travels(arr: pa.Array) -> bool:
    """operator.rs:52-60: whether a column is uploaded rather than kept on
    the host."""
    return device_type(arr.type) and not _null_element_in_valid_row(arr)


def transfer_bytes(batch: pa.RecordBatch) -> int:
    """operator.rs:147-167: the values and validity bytes a transfer of
    `batch` moves (a slice counts its own rows, not its parent's buffers);
    a column no device holds counts nothing."""
    total = 0
    for arr in batch.columns:
        t, n = arr.type, len(arr)
        if not device_type(t):
            continue
        width = t.list_size * 4 if pa.types.is_fixed_size_list(t) else t.bit_width // 8
        total += n * width
        if arr.null_count:
            total += (n + 7) // 8
    return total


def _col(batch: pa.RecordBatch, index: int) -> pa.Array:
    if not 0 <= index < batch.num_columns:
        raise PlanError(f"column index {index} out of range for a batch with {batch.num_columns} columns")
    return batch.column(index)


def _unique(xs) -> list:
    out = []
    for x in xs:
        if x not in out:
            out.append(x)
    return out


@dataclass
class SidePlan:
    """operator.rs:87-124: the columns of one input that travel, in
    sub-batch order, and whether a row-id column follows them."""

    device_cols: list
    row_ids: bool = False

    def position(self, original: int) -> Optional[int]:
        return self.device_cols.index(original) if original in self.device_cols else None

    def row_id_position(self) -> Optional[int]:
        return len(self.device_cols) if self.row_ids else None

    def sub_batch(self, batch: pa.RecordBatch) -> pa.RecordBatch:
        arrays = [_col(batch, c) for c in self.device_cols]
        names = [batch.schema.names[c] for c in self.device_cols]
        if self.row_ids:
            arrays.append(pa.array(np.arange(batch.num_rows, dtype=np.int64)))
            names.append(ROW_ID)
        return pa.RecordBatch.from_arrays(arrays, names=names)

    def remap(self, e):
        """An expression over the sub-batch (operator.rs:126-145)."""
        if e is None:
            return None
        if isinstance(e, plan.Col):
            pos = self.position(e.index)
            if pos is None:
                raise PlanError(f"column {e.index} missing from the device plan")
            return plan.Col(pos)
        if isinstance(e, plan.Bin):
            return plan.Bin(e.kind, self.remap(e.left), self.remap(e.right), e.op)
        return e


def side_plan(key: int, selected, travels_fn) -> SidePlan:
    """operator.rs:809-829: the key, then each selected column that
    travels; any other selected column is gathered on the host."""
    p = SidePlan([key])
    for index in selected:
        if travels_fn(index):
            if index not in p.device_cols:
                p.device_cols.append(index)
        else:
            p.row_ids = True
    return p


def _take(arr: pa.Array, ids: Optional[pa.Array]) -> pa.Array:
    if ids is None:
        raise ExecutionError("row ids missing for host gather")
    return arr.take(ids)


# ------------------------------------------------------------------ the join
class JoinBuild:
    """operator.rs:170-240: the build side of a hash join, shared by every
    probe batch. Which of its columns travel is decided once; its rows count
    into `rows_in` once, with the first probe piece any operator emits."""

    def __init__(self, batch: pa.RecordBatch, key: int):
        _col(batch, key)
        self.batch = batch
        self.key = key
        self.uploadable = [travels(c) for c in batch.columns]
        self._counted = False
        self._lock = threading.Lock()
        self.host_state: dict = {}  # the CPU path's own index, built when first needed

    def column_uploadable(self, index: int) -> bool:
        if not 0 <= index < len(self.uploadable):
            raise PlanError(f"column index {index} out of range for a build side with "
                            f"{len(self.uploadable)} columns")
        return self.uploadable[index]

    def uncounted_rows(self) -> int:
        with self._lock:
            if self._counted:
                return 0
            self._counted = True
            return self.batch.num_rows


@dataclass
class _DeviceBuild:
    plan: SidePlan
    batch: DeviceBatch
    table: object


class JoinPlan:
    """operator.rs:258-371: the columns each side moves and where each output
    column comes from (`selected` indexes `left ++ build`)."""

    def __init__(self, left: pa.RecordBatch, build: JoinBuild, left_key: int,
                 projection: Optional[Sequence[int]]):
        self.width = left.num_columns
        total = self.width + build.batch.num_columns
        self.selected = list(range(total)) if projection is None else list(projection)
        for i in self.selected:
            if not 0 <= i < total:
                raise PlanError(f"join projection index {i} out of range")
        self.left = side_plan(left_key, [i for i in self.selected if i < self.width],
                              lambda i: travels(_col(left, i)))
        self.right = side_plan(build.key, [i - self.width for i in self.selected if i >= self.width],
                               build.column_uploadable)

    def assemble(self, dev_out: pa.RecordBatch, left: pa.RecordBatch, build: pa.RecordBatch) -> pa.RecordBatch:
        lw = len(self.left.device_cols) + int(self.left.row_ids)
        lp, rp = self.left.row_id_position(), self.right.row_id_position()
        left_ids = dev_out.column(lp) if lp is not None else None
        right_ids = dev_out.column(lw + rp) if rp is not None else None
        cols, fields = [], []
        for index in self.selected:
            if index < self.width:
                src, sp, ids, off, local = left, self.left, left_ids, 0, index
            else:
                src, sp, ids, off, local = build, self.right, right_ids, lw, index - self.width
            fields.append(src.schema.field(local))
            pos = sp.position(local)
            cols.append(dev_out.column(off + pos) if pos is not None else _take(src.column(local), ids))
        return pa.RecordBatch.from_arrays(cols, schema=pa.schema(fields))


@dataclass
class _DeviceProbe:
    stream: int
    build: _DeviceBuild
    left: DeviceBatch
    matches: object
    plan: JoinPlan
    next: int = 0


@dataclass
class _HostProbe:
    matches: object
    next: int
    why: Unsupported


@dataclass
class JoinProbe:
    """operator.rs:374-396: one probe batch on its way out in pieces."""

    left: pa.RecordBatch
    left_key: int
    projection: Optional[list]
    batch_size: int
    first: Optional[tuple]  # (start time, rows_in) until the first piece is out
    path: object  # _DeviceProbe | _HostProbe


# ---------------------------------------------------------- the aggregation
class AggregateGroups:
    """The running groups of one aggregation over many batches (the role of
    the reference's `GroupedAggregator` in `aggregate_into`): resident on the
    device while batches run there; the CPU path's aggregator holds the
    batches it ran, and merges the device's groups at `finish`."""

    def __init__(self, op: "GpuOperator", schema_batch: pa.RecordBatch, spec: AggregateSpec):
        self.op = op
        self.schema_batch = schema_batch.slice(0, 0)
        self.spec = spec
        self.state = None  # AggState on the device
        self.host = None  # the CPU path's aggregator
        self._plan: Optional[SidePlan] = None
        self._dev_spec: Optional[AggregateSpec] = None

    def device_plan(self, batch: pa.RecordBatch) -> tuple[SidePlan, AggregateSpec]:
        """operator.rs:1040-1075: only the key and the aggregate inputs (and
        the fused filter's columns) travel."""
        if self._plan is None:
            s = self.spec
            cols = ([s.group_by] if s.group_by >= 0 else [])
            for a in s.aggregates:
                if a.input is not None:
                    cols += plan.columns_of(a.input)
            if s.filter is not None:
                cols += plan.columns_of(s.filter)
            p = SidePlan(_unique(cols))
            dev = AggregateSpec(p.position(s.group_by) if s.group_by >= 0 else -1,
                                [Aggregate(a.func, p.remap(a.input), a.name) for a in s.aggregates],
                                p.remap(s.filter), s.groups_hint)
            self._plan, self._dev_spec = p, dev
        for c in self._plan.device_cols:
            if not travels(_col(batch, c)):
                raise Unsupported("rocm.column_type", f"aggregate input column {c} ({batch.schema.field(c).type}) "
                                  "does not travel to the device")
        return self._plan, self._dev_spec

    def finish(self) -> pa.RecordBatch:
        op = self.op
        start = time.perf_counter()
        if self.state is None and self.host is None:
            # no batch at all: the empty aggregation of the input's schema
            try:
                p, dev_spec = self.device_plan(self.schema_batch)
                if not op.backend.supports_aggregate():
                    raise Unsupported("backend.operation", "the backend does not support aggregation")
                self._create(p.sub_batch(self.schema_batch), dev_spec)
            except Unsupported as why:
                self.host = op._cpu(why).aggregator(self.schema_batch, self.spec)
        partial = None
        if self.state is not None:
            s = op.streams[0]
            out = self.state.finish(s)
            partial = op.backend.download(s, out)
            s.synchronize()
            op._transfer(0, transfer_bytes(partial))
        if self.host is not None:
            if partial is not None:
                self.host.merge(partial)
            result = self.host.finish()
        else:
            result = partial
        with op._lock:
            op.stats.rows_out += result.num_rows
            op.stats.elapsed_s += time.perf_counter() - start
        return result

    def _create(self, sub: pa.RecordBatch, dev_spec: AggregateSpec) -> None:
        op = self.op
        s = op.streams[0]
        like = op.backend.upload(s, sub)
        self.state = op.backend.aggregate_state(s, like, dev_spec)


# -------------------------------------------------------------- the operator
class GpuOperator:
    """operator.rs:399-1145: runs one operator's batches on the backend, with
    per-batch fallback to the caller's CPU path and counters. Two streams are
    used round-robin across batches (operator.rs:480-503); an aggregation's
    resident groups are updated on the first."""

    def __init__(self, backend: RocmBackend, cpu: Optional[CpuPath] = None, streams: int = 2):
        self.backend = backend
        self.cpu = cpu
        self.streams = [backend.create_stream() for _ in range(max(1, int(streams)))]
        self.stats = OperatorStats()
        self._next = 0
        self._lock = threading.Lock()
        self._build_state = None  # ("ready", _DeviceBuild) or ("declined", Unsupported)
        self._warned = False

    # -- plumbing
    def _stream(self) -> int:
        with self._lock:
            i = self._next % len(self.streams)
            self._next += 1
        return i

    def _record(self, rows_in: int, rows_out: int, start: float, moved=None) -> None:
        with self._lock:
            st = self.stats
            st.batches += 1
            st.rows_in += rows_in
            st.rows_out += rows_out
            st.elapsed_s += time.perf_counter() - start
            if moved is not None:
                st.bytes_h2d += moved[0]
                st.bytes_d2h += moved[1]

    def _transfer(self, h2d: int, d2h: int) -> None:
        with self._lock:
            self.stats.bytes_h2d += h2d
            self.stats.bytes_d2h += d2h

    def _cpu(self, why: Unsupported) -> CpuPath:
        if self.cpu is None:
            raise why
        return self.cpu

    def _fell_back(self, why: Unsupported) -> None:
        """operator.rs:547-563: counted per batch, said once per operator."""
        with self._lock:
            self.stats.fallback_batches += 1
            key = getattr(why, "feature", type(why).__name__)
            self.stats.fallbacks[key] = self.stats.fallbacks.get(key, 0) + 1
            first = not self._warned
            self._warned = True
        if first:
            log.warning("operator placed on the device is running the CPU path: %s", why)

    def _no_device_path(self, supported: bool, what: str) -> Optional[Unsupported]:
        return None if supported else Unsupported("backend.operation", f"the backend does not support {what}")

    def _round_trip(self, stream, sub: pa.RecordBatch, kernel) -> tuple[pa.RecordBatch, tuple[int, int]]:
        """operator.rs:565-577: upload, run, download, wait."""
        dev = self.backend.upload(stream, sub)
        out = kernel(stream, dev)
        host = self.backend.download(stream, out)
        stream.synchronize()
        return host, (transfer_bytes(sub), transfer_bytes(host))

    # -- filter + projection (operator.rs:579-705)
    def filter_project(self, batch: pa.RecordBatch, predicate, projection: Sequence,
                       names: Optional[Sequence[str]] = None) -> pa.RecordBatch:
        start = time.perf_counter()
        proj = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        why = self._no_device_path(self.backend.supports_predicate(predicate), "this predicate")
        out = moved = None
        if why is None:
            try:
                out, moved = self._filter_on_device(batch, predicate, proj, names)
            except Unsupported as e:
                why = e
        if why is not None:
            self._fell_back(why)
            out = self._cpu(why).filter_project(batch, predicate, projection, names)
        self._record(batch.num_rows, out.num_rows, start, moved)
        return out

    def _filter_on_device(self, batch, predicate, proj, names):
        # a plain projected column may stay on the host; a predicate or a
        # computed projection reads its columns on the device
        host_only = [isinstance(p, plan.Col) and not travels(_col(batch, p.index)) for p in proj]
        need = plan.columns_of(predicate) if predicate is not None else []
        for p, h in zip(proj, host_only):
            if not h:
                need += plan.columns_of(p)
        for c in need:
            if not travels(_col(batch, c)):
                raise Unsupported("rocm.column_type", f"column {c} ({batch.schema.field(c).type}) is read on the "
                                  "device but does not travel")
        sp = SidePlan(_unique(need), any(host_only))
        sub = sp.sub_batch(batch)
        dev_pred = sp.remap(predicate)
        dev_proj = [sp.remap(p) for p, h in zip(proj, host_only) if not h]
        if sp.row_ids:
            dev_proj.append(plan.Col(sp.row_id_position()))
        dev_out, moved = self._round_trip(
            self.streams[self._stream()], sub,
            lambda s, dev: self.backend.filter_project(s, dev, dev_pred, dev_proj))
        ids = dev_out.column(dev_out.num_columns - 1) if sp.row_ids else None
        cols, fields, k = [], [], 0
        for i, (p, h) in enumerate(zip(proj, host_only)):
            nm = names[i] if names is not None else (batch.schema.names[p.index] if isinstance(p, plan.Col) else f"p{i}")
            if h:
                arr = _take(batch.column(p.index), ids)
            else:
                arr = dev_out.column(k)
                k += 1
            cols.append(arr)
            fields.append(pa.field(nm, arr.type))
        return pa.RecordBatch.from_arrays(cols, schema=pa.schema(fields)), moved

    # -- hash join (operator.rs:705-992)
    def hash_join(self, left: pa.RecordBatch, build: JoinBuild, left_key: int,
                  projection: Optional[Sequence[int]] = None, batch_size: int = 8192) -> JoinProbe:
        """Probes one batch and chooses its path; its output comes from
        `next_join_chunk` in batches of at most `batch_size` rows, as
        `left ++ build` columns or `projection` (indices into that)."""
        start = time.perf_counter()
        why = self._no_device_path(self.backend.supports_hash_join(), "hash joins")
        path = None
        if why is None:
            path = self._probe_on_device(left, build, left_key, projection)
            if isinstance(path, Unsupported):
                why, path = path, None
        if path is None:
            path = _HostProbe(self._cpu(why).join_matches(left, left_key, build), 0, why)
        return JoinProbe(left, left_key, None if projection is None else list(projection),
                         max(1, int(batch_size)), (start, left.num_rows + build.uncounted_rows()), path)

    def next_join_chunk(self, probe: JoinProbe, build: JoinBuild) -> Optional[pa.RecordBatch]:
        """The next piece of the probe batch's output, or None once all of it
        is out. A probe batch always yields at least one (possibly empty)
        piece, which carries its rows in `rows_in`."""
        started = time.perf_counter()
        while True:
            path = probe.path
            total, nxt = len(path.matches), path.next
            remaining = total - nxt
            if remaining == 0 and probe.first is None:
                return None
            rows = min(remaining, probe.batch_size)
            if isinstance(path, _DeviceProbe):
                try:
                    out, moved = self._gather_on_device(path, probe.left, build, rows)
                except Unsupported as why:
                    probe.path = _HostProbe(self._cpu(why).join_matches(probe.left, probe.left_key, build), nxt, why)
                    continue
                path.next += rows
                break
            self._fell_back(path.why)
            out = self.cpu.join_gather(probe.left, build, path.matches, nxt, rows, probe.projection)
            path.next += rows
            moved = None
            break
        start, rows_in = probe.first if probe.first is not None else (started, 0)
        probe.first = None
        self._record(rows_in, out.num_rows, start, moved)
        return out

    def join_all(self, left: pa.RecordBatch, build: JoinBuild, left_key: int,
                 projection: Optional[Sequence[int]] = None, batch_size: int = 8192) -> list[pa.RecordBatch]:
        """Every piece of one probe batch's output."""
        probe = self.hash_join(left, build, left_key, projection, batch_size)
        pieces = []
        while (piece := self.next_join_chunk(probe, build)) is not None:
            pieces.append(piece)
        return pieces

    def _device_build(self, stream, build: JoinBuild, sp: SidePlan) -> _DeviceBuild:
        """operator.rs:831-918: uploaded and indexed on first use, reused by
        every later probe batch; a backend that declines is not asked again."""
        st = self._build_state
        if st is not None:
            if st[0] == "declined":
                raise st[1]
            if st[1].plan == sp:
                return st[1]
        try:
            sub = sp.sub_batch(build.batch)
            dev = self.backend.upload(stream, sub)
            self._transfer(transfer_bytes(sub), 0)
            table = self.backend.prepare_join_build(stream, dev, sp.position(build.key))
            # a later batch may probe on the other stream
            stream.synchronize()
        except Unsupported as why:
            if self._build_state is None:
                self._build_state = ("declined", why)
            raise
        fresh = _DeviceBuild(sp, dev, table)
        if self._build_state is None:
            self._build_state = ("ready", fresh)
        return fresh

    def _probe_on_device(self, left, build, left_key, projection):
        jp = JoinPlan(left, build, left_key, projection)
        sub = jp.left.sub_batch(left)
        key = jp.left.position(left_key)
        i = self._stream()
        on = self.streams[i]
        try:
            dev_build = self._device_build(on, build, jp.right)
            dev_left = self.backend.upload(on, sub)
            self._transfer(transfer_bytes(sub), 0)
            matches = self.backend.probe_join(on, dev_left, key, dev_build.table)
        except Unsupported as why:
            return why
        return _DeviceProbe(i, dev_build, dev_left, matches, jp)

    def _gather_on_device(self, dp: _DeviceProbe, left, build: JoinBuild, rows: int):
        s = self.streams[dp.stream]
        out = self.backend.hash_join(s, dp.left, dp.build.batch, dp.build.table, dp.matches, dp.next, rows)
        host = self.backend.download(s, out)
        s.synchronize()
        return dp.plan.assemble(host, left, build.batch), (0, transfer_bytes(host))

    # -- grouped aggregation (operator.rs:994-1075)
    def new_groups(self, schema_batch: pa.RecordBatch, spec: AggregateSpec) -> AggregateGroups:
        return AggregateGroups(self, schema_batch, spec)

    def aggregate(self, batch: pa.RecordBatch, spec: AggregateSpec) -> pa.RecordBatch:
        """Grouped aggregation of one complete batch."""
        groups = self.new_groups(batch, spec)
        self.aggregate_into(batch, spec, groups)
        return groups.finish()

    def aggregate_into(self, batch: pa.RecordBatch, spec: AggregateSpec, groups: AggregateGroups) -> None:
        """Folds one batch into `groups`: on the device when it runs there,
        else by the CPU path's aggregator. Rows out are counted at finish."""
        start = time.perf_counter()
        why = self._no_device_path(self.backend.supports_aggregate(), "aggregation")
        moved = None
        if why is None:
            try:
                sp, dev_spec = groups.device_plan(batch)
                sub = sp.sub_batch(batch)
                s = self.streams[0]
                dev = self.backend.upload(s, sub)
                if groups.state is None:
                    groups.state = self.backend.aggregate_state(s, dev, dev_spec)
                # no wait: the batch's memory is released in stream order,
                # and the next upload on this stream waits for the update
                groups.state.update(s, dev)
                moved = (transfer_bytes(sub), 0)
            except Unsupported as e:
                why = e
        if why is not None:
            self._fell_back(why)
            if groups.host is None:
                groups.host = self._cpu(why).aggregator(groups.schema_batch, spec)
            groups.host.update(batch)
        self._record(batch.num_rows, 0, start, moved)

    # -- vector distance (operator.rs:1077-1144)
    def vector_distance(self, batch: pa.RecordBatch, column: int, query, metric: DistanceMetric,
                        output_name: str) -> pa.RecordBatch:
        """Appends a nullable Float32 distance column to the batch; only the
        vector column travels."""
        start = time.perf_counter()
        out = moved = None
        try:
            if not travels(_col(batch, column)):
                raise Unsupported("device.vector_column", f"column {column} does not travel to the device")
            sub = SidePlan([column]).sub_batch(batch)
            dev_out, moved = self._round_trip(
                self.streams[self._stream()], sub,
                lambda s, dev: self.backend.vector_distance(s, dev, 0, query, metric, output_name))
            dist = dev_out.column(dev_out.num_columns - 1)
            out = pa.RecordBatch.from_arrays(
                list(batch.columns) + [dist],
                schema=pa.schema(list(batch.schema) + [pa.field(output_name, pa.float32())]))
        except Unsupported as why:
            self._fell_back(why)
            out = self._cpu(why).vector_distance(batch, column, query, metric, output_name)
        self._record(batch.num_rows, out.num_rows, start, moved)
        return out

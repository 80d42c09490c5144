"""Python face of the CPU oracle (oracle/oxide_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

It restates crates/oxidelake-device/src/cpu/kernels.rs over pyarrow
RecordBatches: `filter_project` (kernels.rs:99-152), `GroupedAggregator`
(kernels.rs:478-818), `JoinTable`/`JoinMatches` (kernels.rs:181-472) and
`vector_distance`/`l2`/`cosine` (kernels.rs:840-927). Expressions use the
same postfix encoding as the device library (oxidelake_amd.plan).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional, Sequence

import numpy as np
import pyarrow as pa

from oxidelake_amd import plan
from oxidelake_amd.plan import AggregateFunction, Encoded, ExprT

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboxide_oracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


class _HostCol(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("dim", C.c_int32),
        ("len", C.c_uint64),
        ("values", C.c_void_p),
        ("validity", C.c_void_p),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "oxide_oracle.c").stat().st_mtime:
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        L.ora_total_order_key.restype = C.c_int64
        L.ora_total_order_key.argtypes = [C.c_double]
        L.ora_from_total_order_key.restype = C.c_double
        L.ora_from_total_order_key.argtypes = [C.c_int64]
        L.ora_eval_predicate.argtypes = [P, C.c_int32, C.c_uint64, P, P]
        L.ora_eval_expr.argtypes = [P, C.c_int32, C.c_uint64, P, P, P, C.POINTER(C.c_int32)]
        L.ora_expr_type.argtypes = [P, C.c_int32, P]
        L.ora_expr_type.restype = C.c_int32
        L.ora_agg_new.restype = P
        L.ora_agg_new.argtypes = [C.c_int32, C.c_int32, P, P, C.c_int32]
        L.ora_agg_update.argtypes = [P, C.c_uint64, P, P, P, P]
        L.ora_agg_merge.argtypes = [P, C.c_uint64, P, P, P, P]
        L.ora_agg_groups.restype = C.c_int64
        L.ora_agg_groups.argtypes = [P]
        L.ora_agg_finish.argtypes = [P, P, P, P, P]
        L.ora_agg_free.argtypes = [P]
        L.ora_join_build.restype = P
        L.ora_join_build.argtypes = [C.c_uint64, P, P]
        L.ora_join_distinct.restype = C.c_uint64
        L.ora_join_distinct.argtypes = [P]
        L.ora_join_rows.restype = C.c_uint64
        L.ora_join_rows.argtypes = [P]
        L.ora_join_probe.restype = C.c_uint64
        L.ora_join_probe.argtypes = [P, C.c_uint64, P, P, P, P]
        L.ora_join_pairs.restype = C.c_int
        L.ora_join_pairs.argtypes = [P, C.c_uint64, P, P, C.c_uint64, C.c_uint64, P, P]
        L.ora_join_free.argtypes = [P]
        L.ora_l2.restype = C.c_float
        L.ora_l2.argtypes = [P, P, C.c_int32]
        L.ora_cosine.restype = C.c_float
        L.ora_cosine.argtypes = [P, P, C.c_int32]
        L.ora_vector_distance.argtypes = [P, P, C.c_uint64, C.c_int32, P, C.c_int32, P, P]
        L.ora_topk64.argtypes = [P, P, C.c_uint64, C.c_int32, P, C.c_int32, C.c_int32, P, P]
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------ arrow <-> np
_NP = {plan.INT32: np.int32, plan.INT64: np.int64, plan.FLOAT32: np.float32, plan.FLOAT64: np.float64}


def oxr_type(t: pa.DataType) -> tuple[int, int]:
    if t == pa.int32():
        return plan.INT32, 0
    if t == pa.int64():
        return plan.INT64, 0
    if t == pa.float32():
        return plan.FLOAT32, 0
    if t == pa.float64():
        return plan.FLOAT64, 0
    if pa.types.is_fixed_size_list(t) and t.value_type == pa.float32():
        return plan.F32_VECTOR, t.list_size
    raise TypeError(f"{t} is not an oracle column type")


def valid_bytes(arr: pa.Array) -> Optional[np.ndarray]:
    """Validity as one byte per row (None when the array has no NULLs)."""
    if arr.null_count == 0:
        return None
    return np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=np.uint8)


def values_np(arr: pa.Array) -> np.ndarray:
    """Values (NULL slots hold whatever the buffer holds) as a contiguous array."""
    t, dim = oxr_type(arr.type)
    if t == plan.F32_VECTOR:
        flat = arr.values  # child, sliced by the list's offset already
        off = arr.offset * dim
        child = np.frombuffer(flat.buffers()[1], dtype=np.float32)
        start = flat.offset + off
        return np.ascontiguousarray(child[start:start + len(arr) * dim]).reshape(len(arr), dim)
    buf = arr.buffers()[1]
    a = np.frombuffer(buf, dtype=_NP[t])
    return np.ascontiguousarray(a[arr.offset:arr.offset + len(arr)])


class HostBatch:
    """A RecordBatch as oracle columns (values + LSB-first validity bitmaps)."""

    def __init__(self, batch: pa.RecordBatch):
        self.batch = batch
        self.n = batch.num_rows
        self.keep = []
        self.cols = (_HostCol * max(1, batch.num_columns))()
        for i, arr in enumerate(batch.columns):
            try:
                t, dim = oxr_type(arr.type)
            except TypeError:
                self.cols[i].type = 0
                continue
            v = values_np(arr)
            vb = valid_bytes(arr)
            bits = None if vb is None else np.packbits(vb, bitorder="little")
            self.keep += [v, bits]
            self.cols[i].type = t
            self.cols[i].dim = dim
            self.cols[i].len = self.n
            self.cols[i].values = v.ctypes.data
            self.cols[i].validity = None if bits is None else bits.ctypes.data


# ------------------------------------------------------------------ filter
def predicate_mask(batch: pa.RecordBatch, predicate: ExprT | None) -> np.ndarray:
    """kernels.rs:99-141 `evaluate_predicate`: True where the row is kept."""
    hb = HostBatch(batch)
    mask = np.zeros(max(1, hb.n), dtype=np.uint8)
    enc = Encoded(predicate)
    rc = lib().ora_eval_predicate(hb.cols, batch.num_columns, hb.n, enc.ref(), _ptr(mask))
    if rc == 1 or rc == 2:
        raise ZeroDivisionError("Divide by zero error")
    if rc != 0:
        raise ValueError("malformed predicate")
    return mask[: hb.n].astype(bool)


def eval_expr(batch: pa.RecordBatch, e: ExprT) -> pa.Array:
    """One scalar expression at every row, as an Arrow array."""
    hb = HostBatch(batch)
    if isinstance(e, plan.Col):
        return batch.column(e.index)
    enc = Encoded(e)
    t = lib().ora_expr_type(hb.cols, batch.num_columns, enc.ref())
    if t <= 0:
        raise ValueError("not a value expression")
    vals = np.zeros(max(1, hb.n), dtype=_NP[t])
    valid = np.zeros(max(1, hb.n), dtype=np.uint8)
    out_t = C.c_int32()
    rc = lib().ora_eval_expr(hb.cols, batch.num_columns, hb.n, enc.ref(), _ptr(vals), _ptr(valid),
                             C.byref(out_t))
    if rc == 1 or rc == 2:
        raise ZeroDivisionError("Divide by zero error")
    if rc != 0:
        raise ValueError("malformed expression")
    vals, valid = vals[: hb.n], valid[: hb.n].astype(bool)
    arrow_t = {plan.INT32: pa.int32(), plan.INT64: pa.int64(), plan.FLOAT32: pa.float32(),
               plan.FLOAT64: pa.float64()}[t]
    return pa.array(vals, type=arrow_t, mask=~valid if not valid.all() else None)


def filter_project(batch: pa.RecordBatch, predicate: ExprT | None, projection: Sequence,
                   names: Optional[Sequence[str]] = None) -> pa.RecordBatch:
    """kernels.rs:144-152: mask, then `filter_record_batch` of the projection.
    Arithmetic projections (beyond the reference's column-index projections)
    are evaluated over the kept rows only, as DataFusion's ProjectionExec
    above a FilterExec: a zero divisor in a dropped row is no error."""
    exprs = [plan.Col(p) if isinstance(p, int) else p for p in projection]
    mask = predicate_mask(batch, predicate)
    idx = np.nonzero(mask)[0]
    kept = batch.take(pa.array(idx, type=pa.int64()))
    cols, fields = [], []
    for i, e in enumerate(exprs):
        arr = eval_expr(kept, e)
        cols.append(arr)
        nm = names[i] if names else (batch.schema.field(e.index).name if isinstance(e, plan.Col) else f"p{i}")
        fields.append(pa.field(nm, arr.type))
    return pa.RecordBatch.from_arrays(cols, schema=pa.schema(fields))


# --------------------------------------------------------------- aggregate
def _agg_input_type(batch: pa.RecordBatch, a: plan.Aggregate) -> int:
    if a.func == AggregateFunction.CountStar:
        return plan.INT64
    hb = HostBatch(batch)
    if isinstance(a.input, plan.Col):
        return oxr_type(batch.schema.field(a.input.index).type)[0]
    return lib().ora_expr_type(hb.cols, batch.num_columns, Encoded(a.input).ref())


class GroupedAggregator:
    """kernels.rs:655-811: streaming grouped aggregation, one key column
    (Int64 or Int32) or none (group_by = -1: one global group)."""

    def __init__(self, schema_batch: pa.RecordBatch, spec: plan.AggregateSpec, sum_mode: int = 0):
        self.spec = spec
        self.key_type = (oxr_type(schema_batch.schema.field(spec.group_by).type)[0]
                         if spec.group_by >= 0 else plan.INT64)
        self.in_types = [_agg_input_type(schema_batch, a) for a in spec.aggregates]
        funcs = np.array([int(a.func) for a in spec.aggregates], dtype=np.int32)
        types = np.array(self.in_types, dtype=np.int32)
        self.h = lib().ora_agg_new(self.key_type, len(funcs), _ptr(funcs), _ptr(types), sum_mode)
        self.names = []
        for a in spec.aggregates:
            if a.name:
                self.names.append(a.name)
            elif a.func == AggregateFunction.CountStar:
                self.names.append("COUNT(*)")
            elif isinstance(a.input, plan.Col):
                self.names.append(f"{a.func.sql}({schema_batch.schema.field(a.input.index).name})")
            else:
                self.names.append(f"{a.func.sql}(expr)")
        self.key_name = schema_batch.schema.field(spec.group_by).name if spec.group_by >= 0 else "group"

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_agg_free(self.h)
            self.h = None

    def _key(self, batch: pa.RecordBatch):
        n = batch.num_rows
        if self.spec.group_by < 0:
            return np.zeros(max(1, n), dtype=np.int64), None
        arr = batch.column(self.spec.group_by)
        keys = values_np(arr).astype(np.int64)
        return keys, valid_bytes(arr)

    def update(self, batch: pa.RecordBatch) -> None:
        """kernels.rs:765-772 (with the spec's filter applied first)."""
        if self.spec.filter is not None:
            mask = predicate_mask(batch, self.spec.filter)
            batch = batch.filter(pa.array(mask))
        n = batch.num_rows
        keys, kvalid = self._key(batch)
        vals, valids, keep = [], [], []
        for a, t in zip(self.spec.aggregates, self.in_types):
            if a.func == AggregateFunction.CountStar:
                v, vb = np.zeros(max(1, n), dtype=np.int64), None
            else:
                arr = eval_expr(batch, a.input)
                v = values_np(arr) if n else np.zeros(1, dtype=_NP[t])
                if t in (plan.INT32, plan.INT64):
                    v = v.astype(np.int64)
                v = np.ascontiguousarray(v)
                vb = valid_bytes(arr)
            keep += [v, vb]
            vals.append(v.ctypes.data)
            valids.append(None if vb is None else vb.ctypes.data)
        varr = (C.c_void_p * max(1, len(vals)))(*vals)
        vvarr = (C.c_void_p * max(1, len(valids)))(*valids)
        lib().ora_agg_update(self.h, n, _ptr(keys), _ptr(kvalid), varr, vvarr)

    def merge(self, partial: pa.RecordBatch) -> None:
        """kernels.rs:776-790: folds a partial laid out [key, agg...]."""
        if partial.num_columns != 1 + len(self.spec.aggregates):
            raise ValueError("a partial aggregate is not laid out as this aggregation's output")
        n = partial.num_rows
        karr = partial.column(0)
        keys = values_np(karr).astype(np.int64) if n else np.zeros(1, dtype=np.int64)
        kvalid = valid_bytes(karr)
        vals, valids, keep = [], [], []
        for i, _ in enumerate(self.spec.aggregates):
            arr = partial.column(1 + i)
            v = values_np(arr) if n else np.zeros(1, dtype=np.int64)
            if v.dtype in (np.int32,):
                v = v.astype(np.int64)
            v = np.ascontiguousarray(v)
            vb = valid_bytes(arr)
            keep += [v, vb]
            vals.append(v.ctypes.data)
            valids.append(None if vb is None else vb.ctypes.data)
        varr = (C.c_void_p * max(1, len(vals)))(*vals)
        vvarr = (C.c_void_p * max(1, len(valids)))(*valids)
        lib().ora_agg_merge(self.h, n, _ptr(keys), _ptr(kvalid), varr, vvarr)

    def groups(self) -> int:
        return int(lib().ora_agg_groups(self.h))

    def output_types(self) -> list[int]:
        out = []
        for a, t in zip(self.spec.aggregates, self.in_types):
            if a.func in (AggregateFunction.Count, AggregateFunction.CountStar):
                out.append(plan.INT64)
            elif a.func == AggregateFunction.Sum:
                out.append(plan.INT64 if t in (plan.INT32, plan.INT64) else plan.FLOAT64)
            else:
                out.append(t)
        return out

    def finish(self) -> pa.RecordBatch:
        """kernels.rs:793-811: groups sorted by key, NULL last."""
        g = self.groups()
        keys = np.zeros(max(1, g), dtype=np.int64)
        kvalid = np.zeros(max(1, g), dtype=np.uint8)
        outs, ovalid = [], []
        otypes = self.output_types()
        for a, ot, it in zip(self.spec.aggregates, otypes, self.in_types):
            dt = np.int64 if ot in (plan.INT64,) or it in (plan.INT32, plan.INT64) else (
                np.float64 if ot == plan.FLOAT64 else np.float32)
            if ot == plan.INT32:
                dt = np.int64
            outs.append(np.zeros(max(1, g), dtype=dt))
            ovalid.append(np.zeros(max(1, g), dtype=np.uint8))
        oarr = (C.c_void_p * max(1, len(outs)))(*[o.ctypes.data for o in outs])
        varr = (C.c_void_p * max(1, len(outs)))(*[v.ctypes.data for v in ovalid])
        lib().ora_agg_finish(self.h, _ptr(keys), _ptr(kvalid), oarr, varr)
        key_t = pa.int32() if self.key_type == plan.INT32 else pa.int64()
        cols, fields = [], []
        if self.spec.group_by >= 0:
            cols.append(pa.array(keys[:g].astype(np.int32 if self.key_type == plan.INT32 else np.int64),
                                 type=key_t, mask=~kvalid[:g].astype(bool)))
            fields.append(pa.field(self.key_name, key_t))
        elif g == 0:
            # SQL: a global aggregate over no rows is one row (COUNT 0, the rest NULL)
            g = 1
            for a, o, v in zip(self.spec.aggregates, outs, ovalid):
                o[0] = 0
                v[0] = a.func in (AggregateFunction.Count, AggregateFunction.CountStar)
        for o, v, ot, nm in zip(outs, ovalid, otypes, self.names):
            at = {plan.INT32: pa.int32(), plan.INT64: pa.int64(), plan.FLOAT32: pa.float32(),
                  plan.FLOAT64: pa.float64()}[ot]
            vals = o[:g].astype(np.int32) if ot == plan.INT32 else o[:g]
            cols.append(pa.array(vals, type=at, mask=~v[:g].astype(bool)))
            fields.append(pa.field(nm, at))
        return pa.RecordBatch.from_arrays(cols, schema=pa.schema(fields))


def aggregate(batch: pa.RecordBatch, spec: plan.AggregateSpec, sum_mode: int = 0) -> pa.RecordBatch:
    """kernels.rs:814-818: one pass."""
    g = GroupedAggregator(batch, spec, sum_mode)
    g.update(batch)
    return g.finish()


# -------------------------------------------------------------------- join
class JoinTable:
    """kernels.rs:181-266 `JoinTable`."""

    def __init__(self, build: pa.RecordBatch, key: int):
        arr = build.column(key)
        self.keys = values_np(arr).astype(np.int64) if build.num_rows else np.zeros(1, np.int64)
        self.valid = valid_bytes(arr)
        self.h = lib().ora_join_build(build.num_rows, _ptr(self.keys), _ptr(self.valid))
        self.build_rows = build.num_rows

    def distinct_keys(self) -> int:
        return int(lib().ora_join_distinct(self.h))

    def rows(self) -> int:
        return int(lib().ora_join_rows(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_join_free(self.h)
            self.h = None


class JoinMatches:
    """kernels.rs:294-407 `JoinMatches`."""

    def __init__(self, probe: pa.RecordBatch, key: int, table: JoinTable):
        n = probe.num_rows
        arr = probe.column(key)
        keys = values_np(arr).astype(np.int64) if n else np.zeros(1, np.int64)
        valid = valid_bytes(arr)
        self.offsets = np.zeros(n + 1, dtype=np.uint64)
        self.starts = np.zeros(max(1, n), dtype=np.uint32)
        self.total = int(lib().ora_join_probe(table.h, n, _ptr(keys), _ptr(valid),
                                              _ptr(self.offsets), _ptr(self.starts)))
        self.n = n
        self.table = table

    def __len__(self) -> int:
        return self.total

    def pairs(self, first: int, rows: int) -> tuple[np.ndarray, np.ndarray]:
        left = np.zeros(max(1, rows), dtype=np.uint32)
        right = np.zeros(max(1, rows), dtype=np.uint32)
        rc = lib().ora_join_pairs(self.table.h, self.n, _ptr(self.offsets), _ptr(self.starts),
                                  first, rows, _ptr(left), _ptr(right))
        if rc != 0:
            raise IndexError(f"matches {first}..+{rows} are outside a probe with {self.total} matches")
        return left[:rows], right[:rows]


def hash_join_gather(left: pa.RecordBatch, right: pa.RecordBatch, matches: JoinMatches,
                     first: int, rows: int, projection: Optional[Sequence[int]] = None) -> pa.RecordBatch:
    """kernels.rs:414-472."""
    li, ri = matches.pairs(first, rows)
    fields = list(left.schema) + list(right.schema)
    width = left.num_columns
    sel = list(range(len(fields))) if projection is None else list(projection)
    cols = []
    for i in sel:
        if i < width:
            cols.append(left.column(i).take(pa.array(li, type=pa.uint32())))
        else:
            cols.append(right.column(i - width).take(pa.array(ri, type=pa.uint32())))
    return pa.RecordBatch.from_arrays(cols, schema=pa.schema([fields[i] for i in sel]))


def hash_join(left: pa.RecordBatch, right: pa.RecordBatch, left_key: int, right_key: int,
              projection: Optional[Sequence[int]] = None) -> pa.RecordBatch:
    """kernels.rs:270-300: build over `right`, probe `left` once."""
    t = JoinTable(right, right_key)
    m = JoinMatches(left, left_key, t)
    return hash_join_gather(left, right, m, 0, len(m), projection)


# ------------------------------------------------------------------ vector
def l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().ora_l2(_ptr(a), _ptr(b), len(a)))


def cosine(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib().ora_cosine(_ptr(a), _ptr(b), len(a)))


def vector_distance(batch: pa.RecordBatch, column: int, query: Sequence[float],
                    metric: plan.DistanceMetric, output_name: str) -> pa.RecordBatch:
    """kernels.rs:841-905: appends a nullable Float32 distance column."""
    arr = batch.column(column)
    t, dim = oxr_type(arr.type)
    if t != plan.F32_VECTOR:
        raise TypeError("expected FixedSizeList<Float32>")
    q = np.ascontiguousarray(np.asarray(query, dtype=np.float32))
    if len(q) != dim:
        raise ValueError(f"query vector has {len(q)} dimensions but the column has {dim}")
    n = batch.num_rows
    vecs = values_np(arr) if n else np.zeros((1, dim), np.float32)
    valid = valid_bytes(arr)
    out = np.zeros(max(1, n), dtype=np.float32)
    vout = np.zeros(max(1, n), dtype=np.uint8)
    lib().ora_vector_distance(_ptr(vecs), _ptr(valid), n, dim, _ptr(q), int(metric), _ptr(out), _ptr(vout))
    d = pa.array(out[:n], type=pa.float32(), mask=~vout[:n].astype(bool))
    return pa.RecordBatch.from_arrays(list(batch.columns) + [d],
                                      schema=pa.schema(list(batch.schema) + [pa.field(output_name, pa.float32())]))


def topk64(vecs: np.ndarray, valid: Optional[np.ndarray], query: np.ndarray,
           metric: plan.DistanceMetric, k: int) -> tuple[np.ndarray, np.ndarray]:
    """Exact top-k by FP64 distance (ties by row)."""
    vecs = np.ascontiguousarray(vecs, dtype=np.float32)
    q = np.ascontiguousarray(query, dtype=np.float32)
    rows = np.zeros(k, dtype=np.int64)
    dist = np.zeros(k, dtype=np.float64)
    v = None if valid is None else np.ascontiguousarray(valid, dtype=np.uint8)
    lib().ora_topk64(_ptr(vecs), _ptr(v), vecs.shape[0], vecs.shape[1], _ptr(q), int(metric), k,
                     _ptr(rows), _ptr(dist))
    return rows, dist


def total_order_key(x: float) -> int:
    return int(lib().ora_total_order_key(float(x)))


def from_total_order_key(k: int) -> float:
    return float(lib().ora_from_total_order_key(int(k)))

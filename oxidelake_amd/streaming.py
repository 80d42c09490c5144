"""The plain drop-in path over host batches: the shape of the reference's
operator (crates/oxidelake-compute/src/operator.rs:589-705): each host
Arrow batch is uploaded, filtered and projected on the device, and the
result downloaded, batch by batch.

With `lanes` > 1, that many host threads each own a HIP stream and take
every lanes-th batch, so one batch's transfers overlap another's kernel and
transfers (the entry points release the interpreter lock while they wait).
Results come back in batch order. A batch the device declines raises
`Unsupported`, as the trait's decline does; the caller runs its CPU path.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from typing import Optional, Sequence

import numpy as np
import pyarrow as pa

from . import plan
from .backend import RocmBackend


class StreamingFilter:
    def __init__(self, be: RocmBackend, predicate, projection: Sequence,
                 names: Optional[Sequence[str]] = None, lanes: int = 1):
        self.be = be
        self.prepared = be.prepare_filter(predicate, projection, names)
        self.lanes = max(1, int(lanes))
        self.streams = [be.create_stream() for _ in range(self.lanes)]

    def run(self, batches: Sequence[pa.RecordBatch]) -> list[pa.RecordBatch]:
        out: list = [None] * len(batches)

        def lane(k: int) -> None:
            s = self.streams[k]
            for i in range(k, len(batches), self.lanes):
                dev = self.be.upload(s, batches[i])
                res = self.prepared.run(s, dev)
                out[i] = self.be.download(s, res)
                del dev, res

        if self.lanes == 1:
            lane(0)
        else:
            with ThreadPoolExecutor(self.lanes) as pool:
                for f in [pool.submit(lane, k) for k in range(self.lanes)]:
                    f.result()
        return out


def remap(e, mapping: dict):
    """The expression over a sub-batch: Col(i) becomes Col(mapping[i])."""
    if isinstance(e, plan.Col):
        return plan.Col(mapping[e.index])
    if isinstance(e, plan.Bin):
        return plan.Bin(e.kind, remap(e.left, mapping), remap(e.right, mapping), e.op)
    return e


class FilterOperator:
    """The caller side of filter_project over host batches, as the
    reference's operator runs it (crates/oxidelake-compute/src/operator.rs:
    603-705): only the columns the predicate and the computed projections
    read travel to the device (SidePlan, remap_predicate); a projected
    column of a type the device does not take stays on the host and is
    gathered there by row id (a row-id column travels instead); a batch
    whose predicate or computed projections read such a column, or that
    the backend declines (Unsupported), runs `cpu_fallback` and is counted.
    """

    def __init__(self, be: RocmBackend, predicate, projection: Sequence, names: Optional[Sequence[str]] = None,
                 cpu_fallback=None):
        self.be = be
        self.predicate = predicate
        self.projection = [plan.Col(p) if isinstance(p, int) else p for p in projection]
        self.names = names
        self.cpu_fallback = cpu_fallback
        self.stream = be.create_stream()
        self.device_batches = 0
        self.fallback_batches = 0
        self._prepared: dict = {}

    def _fallback(self, batch: pa.RecordBatch, why: Exception):
        self.fallback_batches += 1
        if self.cpu_fallback is None:
            raise why
        return self.cpu_fallback(batch)

    def process(self, batch: pa.RecordBatch) -> pa.RecordBatch:
        from .backend import Unsupported
        from .loader import device_type

        ok = [device_type(f.type) for f in batch.schema]
        need = set(plan.columns_of(self.predicate)) if self.predicate is not None else set()
        host_cols = []  # projected plain columns that stay on the host
        for p in self.projection:
            if isinstance(p, plan.Col) and not ok[p.index]:
                host_cols.append(p.index)
            else:
                need |= set(plan.columns_of(p))
        if any(not ok[i] for i in need):
            return self._fallback(batch, Unsupported("rocm.types", "a filter or computed column the device does not take"))
        travel = sorted(need | {p.index for p in self.projection if isinstance(p, plan.Col) and ok[p.index]})
        mapping = {c: k for k, c in enumerate(travel)}
        arrays = [batch.column(c) for c in travel]
        names = [batch.schema.names[c] for c in travel]
        if host_cols:
            arrays.append(pa.array(np.arange(batch.num_rows, dtype=np.int64)))
            names.append("__row")
        sub = pa.RecordBatch.from_arrays(arrays, names=names)
        key = (tuple(travel), tuple(ok))
        if key not in self._prepared:
            dev_proj = [remap(p, mapping) for p in self.projection if not (isinstance(p, plan.Col) and not ok[p.index])]
            if host_cols:
                dev_proj.append(plan.Col(len(travel)))
            pred = remap(self.predicate, mapping) if self.predicate is not None else None
            self._prepared[key] = self.be.prepare_filter(pred, dev_proj)
        try:
            dev = self.be.upload(self.stream, sub)
            out = self.be.download(self.stream, self._prepared[key].run(self.stream, dev))
        except Unsupported as e:
            return self._fallback(batch, e)
        self.device_batches += 1
        rows = out.column(out.num_columns - 1) if host_cols else None
        cols, k = [], 0
        for p in self.projection:
            if isinstance(p, plan.Col) and not ok[p.index]:
                cols.append(batch.column(p.index).take(rows))
            else:
                cols.append(out.column(k))
                k += 1
        names = self.names or [batch.schema.names[p.index] if isinstance(p, plan.Col) else f"p{i}"
                               for i, p in enumerate(self.projection)]
        return pa.RecordBatch.from_arrays(cols, names=list(names))

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

import pyarrow as pa

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


class _CallableCpu:
    def __init__(self, fn):
        self.fn = fn

    def filter_project(self, batch, predicate, projection, names):
        return self.fn(batch)


class FilterOperator:
    """One filter + projection over host batches through `GpuOperator`
    (operator.py): the sub-batch cut, host-side columns gathered by row id,
    and `cpu_fallback(batch)` for a batch the device declines, counted."""

    def __init__(self, be: RocmBackend, predicate, projection: Sequence, names: Optional[Sequence[str]] = None,
                 cpu_fallback=None):
        from .operator import GpuOperator

        self.predicate = predicate
        self.projection = list(projection)
        self.names = names
        self.op = GpuOperator(be, _CallableCpu(cpu_fallback) if cpu_fallback is not None else None, streams=1)

    @property
    def device_batches(self) -> int:
        return self.op.stats.batches - self.op.stats.fallback_batches

    @property
    def fallback_batches(self) -> int:
        return self.op.stats.fallback_batches

    def process(self, batch: pa.RecordBatch) -> pa.RecordBatch:
        return self.op.filter_project(batch, self.predicate, self.projection, self.names)

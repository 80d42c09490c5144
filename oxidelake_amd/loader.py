"""Parquet to HBM: a file decoded on the host (pyarrow) and made resident on
the device as one contiguous batch, the form the operators scan in one
launch (filter_project_batches cuts it into batches again).

The reference reads Parquet on the host and uploads each batch as the
operator meets it (crates/oxidelake-compute/src/operator.rs:589-600); GPU
Parquet decode is out of its scope (docs/architecture.md:112) and of this
one. Types the device takes: Int32, Int64, Float32, Float64 and
FixedSizeList<Float32> (crates/oxidelake-device/src/columns.rs:25-309);
other columns are left out when `columns` is None, refused when named.
"""
from __future__ import annotations

from typing import Optional, Sequence

import pyarrow as pa
import pyarrow.parquet as pq

from .backend import DeviceBatch, FormatError, RocmBackend, Stream

_DEVICE_TYPES = (pa.int32(), pa.int64(), pa.float32(), pa.float64())


def device_type(t: pa.DataType) -> bool:
    if t in _DEVICE_TYPES:
        return True
    return pa.types.is_fixed_size_list(t) and t.value_type == pa.float32()


def load_parquet(be: RocmBackend, stream: Stream, path: str,
                 columns: Optional[Sequence[str]] = None) -> DeviceBatch:
    """Reads `path` (every row group) into one HBM-resident batch."""
    pf = pq.ParquetFile(path)
    schema = pf.schema_arrow
    if columns is None:
        columns = [f.name for f in schema if device_type(f.type)]
    else:
        for c in columns:
            if not device_type(schema.field(c).type):
                raise FormatError(f"column {c!r} has type {schema.field(c).type}, which the device does not take")
    table = pf.read(columns=list(columns)).combine_chunks()
    batches = table.to_batches(max_chunksize=max(1, table.num_rows))
    batch = batches[0] if batches else pa.RecordBatch.from_pylist([], schema=table.schema)
    return be.upload(stream, batch)


def write_parquet(batch: pa.RecordBatch | pa.Table, path: str, row_group_rows: int = 8192) -> None:
    """Writes a table the way the reference's generator does (row groups of
    `row_group_rows`, 8,192 by default)."""
    table = batch if isinstance(batch, pa.Table) else pa.Table.from_batches([batch])
    pq.write_table(table, path, row_group_size=row_group_rows)

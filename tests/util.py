"""Comparison helpers for parity tests."""
from __future__ import annotations

import numpy as np
import pyarrow as pa


def _valid(arr: pa.Array) -> np.ndarray:
    if arr.null_count == 0:
        return np.ones(len(arr), dtype=bool)
    return np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=bool)


def _bits(arr: pa.Array) -> np.ndarray:
    """Values as raw bits (so NaN payloads and signed zeros compare exactly)."""
    t = arr.type
    if pa.types.is_fixed_size_list(t):
        flat = arr.flatten() if arr.null_count == 0 else None
        if flat is None:
            # flatten() drops NULL rows; read the child directly
            dim = t.list_size
            child = np.frombuffer(arr.values.buffers()[1], dtype=np.uint32)
            s = arr.values.offset + arr.offset * dim
            return child[s:s + len(arr) * dim].reshape(len(arr), dim)
        return np.frombuffer(flat.buffers()[1], dtype=np.uint32)[flat.offset:flat.offset + len(flat)].reshape(len(arr), t.list_size)
    width = {pa.int32(): np.uint32, pa.float32(): np.uint32, pa.int64(): np.uint64, pa.float64(): np.uint64}[t]
    buf = np.frombuffer(arr.buffers()[1], dtype=width)
    return buf[arr.offset:arr.offset + len(arr)]


def assert_arrays_equal(got: pa.Array, want: pa.Array, what: str = "") -> None:
    assert got.type == want.type, f"{what}: type {got.type} != {want.type}"
    assert len(got) == len(want), f"{what}: {len(got)} rows != {len(want)}"
    if not (pa.types.is_fixed_size_list(got.type) or got.type in (pa.int32(), pa.int64(), pa.float32(),
                                                                   pa.float64())):
        # a column the device does not hold (gathered on the host): by value
        assert got.to_pylist() == want.to_pylist(), f"{what}: values differ"
        return
    gv, wv = _valid(got), _valid(want)
    bad = np.nonzero(gv != wv)[0]
    assert bad.size == 0, f"{what}: validity differs at rows {bad[:10]}"
    gb, wb = _bits(got), _bits(want)
    if gb.ndim == 2:
        diff = np.nonzero((gb != wb).any(axis=1) & wv)[0]
    else:
        diff = np.nonzero((gb != wb) & wv)[0]
    assert diff.size == 0, (
        f"{what}: {diff.size} values differ, first rows {diff[:5]}: "
        f"got {got.take(pa.array(diff[:5])).to_pylist()} want {want.take(pa.array(diff[:5])).to_pylist()}"
    )


def assert_batches_equal(got: pa.RecordBatch, want: pa.RecordBatch, names: bool = True) -> None:
    assert got.num_columns == want.num_columns, f"{got.num_columns} columns != {want.num_columns}"
    assert got.num_rows == want.num_rows, f"{got.num_rows} rows != {want.num_rows}"
    for i in range(got.num_columns):
        if names:
            assert got.schema.field(i).name == want.schema.field(i).name
        assert_arrays_equal(got.column(i), want.column(i), f"column {i} ({want.schema.field(i).name})")


def assert_close_rel(got: np.ndarray, want: np.ndarray, rel: float, what: str = "") -> None:
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    scale = np.maximum(np.abs(want), np.finfo(np.float64).tiny)
    err = np.abs(got - want) / scale
    worst = int(np.argmax(err)) if err.size else 0
    assert err.size == 0 or err[worst] <= rel, (
        f"{what}: relative error {err[worst]:.3e} > {rel:.1e} at {worst}: got {got[worst]!r} want {want[worst]!r}"
    )

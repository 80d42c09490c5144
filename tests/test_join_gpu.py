"""GPU parity: hash join build / probe / gather vs the CPU oracle.

Bit-exact, including the match order (probe row, then build order within a
key), which the reference relies on to finish a probe batch on the CPU from
any match (crates/oxidelake-device/src/backend.rs:148-152).
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd.backend import ExecutionError
from tests.util import assert_batches_equal

pytestmark = pytest.mark.gpu


def device_join(backend, stream, left, right, lk, rk, projection=None, chunk=None):
    dl, dr = backend.upload(stream, left), backend.upload(stream, right)
    table = backend.prepare_join_build(stream, dr, rk)
    m = backend.probe_join(stream, dl, lk, table)
    if chunk is None:
        return backend.download(stream, backend.hash_join(stream, dl, dr, table, m, 0, len(m), projection)), table, m
    parts, first = [], 0
    while first < len(m):
        rows = min(chunk, len(m) - first)
        parts.append(backend.download(stream, backend.hash_join(stream, dl, dr, table, m, first, rows, projection)))
        first += rows
    return parts, table, m


def ref_left():
    return pa.record_batch({"k": pa.array([1, 2, None, 2, 5], pa.int64()),
                            "v": pa.array([0.5, None, 2.5, 3.5, -1.0], pa.float64())})


def test_join_skips_null_keys_and_multiplies_matches(backend, stream, oracle):
    # kernels.rs:1033-1058
    right = pa.record_batch({"rk": pa.array([2, 2, None, 9], pa.int64()),
                             "x": pa.array([10, 11, 12, 13], pa.int64())})
    got, table, m = device_join(backend, stream, ref_left(), right, 0, 0)
    assert got.num_rows == 4 and got.num_columns == 4
    assert sorted(got.column(3).to_pylist()) == [10, 10, 11, 11]
    assert_batches_equal(got, oracle.hash_join(ref_left(), right, 0, 0))


def test_chunked_gather_equals_one_gather(backend, stream, oracle):
    # kernels.rs:1063-1121: every chunk size 1..8
    right = pa.record_batch({"rk": pa.array([2, 2, None, 9, 2, 5], pa.int64()),
                             "x": pa.array([1, 2, 3, 4, 5, 6], pa.int64())})
    left = ref_left()
    whole = oracle.hash_join(left, right, 0, 0)
    proj = [0, 1, 3]
    for chunk in range(1, 9):
        parts, table, m = device_join(backend, stream, left, right, 0, 0, proj, chunk)
        assert len(m) == 7
        got = pa.Table.from_batches(parts).combine_chunks().to_batches()[0]
        assert_batches_equal(got, whole.select(proj))
    dl, dr = backend.upload(stream, left), backend.upload(stream, right)
    t = backend.prepare_join_build(stream, dr, 0)
    m = backend.probe_join(stream, dl, 0, t)
    with pytest.raises(ExecutionError):
        backend.hash_join(stream, dl, dr, t, m, 7, 1)
    other = backend.prepare_join_build(stream, dr, 0)
    with pytest.raises(ExecutionError):
        backend.hash_join(stream, dl, dr, other, m, 0, 1)


@pytest.mark.parametrize("n_l,n_r,keys", [(0, 5, 5), (7, 0, 5), (64, 33, 20), (1000, 501, 10),
                                          (5000, 3000, 100000), (20000, 20000, 2**40)])
@pytest.mark.parametrize("seed", [1, 7])
def test_random_parity(backend, stream, oracle, n_l, n_r, keys, seed):
    rng = np.random.default_rng(seed * 31 + n_l)
    lk = rng.integers(-keys // 2, keys, n_l).astype(np.int64)
    rk = rng.integers(-keys // 2, keys, n_r).astype(np.int64)
    left = pa.record_batch({"k": pa.array(lk, mask=rng.random(n_l) < 0.1),
                            "a": pa.array(rng.normal(size=n_l)),
                            "b": pa.array(rng.integers(0, 9, n_l).astype(np.int32))})
    right = pa.record_batch({"k": pa.array(rk, mask=rng.random(n_r) < 0.1),
                             "c": pa.array(rng.normal(size=n_r).astype(np.float32), mask=rng.random(n_r) < 0.2),
                             "e": pa.FixedSizeListArray.from_arrays(
                                 pa.array(rng.normal(size=n_r * 3).astype(np.float32)), 3)})
    got, table, m = device_join(backend, stream, left, right, 0, 0)
    want = oracle.hash_join(left, right, 0, 0)
    assert_batches_equal(got, want)
    d, r, dense = table.info()
    ot = oracle.JoinTable(right, 0)
    assert (d, r) == (ot.distinct_keys(), ot.rows())


def test_unique_dense_and_hashed_keys(backend, stream, oracle):
    rng = np.random.default_rng(3)
    n = 100_000
    pk = rng.permutation(n).astype(np.int64)  # dense: a permutation of 0..n
    right = pa.record_batch({"pk": pa.array(pk), "grp": pa.array(rng.integers(0, 1000, n).astype(np.int32))})
    fk = rng.integers(-10, n + 10, 300_000).astype(np.int64)
    left = pa.record_batch({"fk": pa.array(fk), "q": pa.array(rng.integers(1, 100, 300_000).astype(np.int32))})
    got, table, m = device_join(backend, stream, left, right, 0, 0)
    assert table.info()[2]  # dense
    assert_batches_equal(got, oracle.hash_join(left, right, 0, 0))
    right2 = pa.record_batch({"pk": pa.array(pk * 1_000_003), "grp": right.column(1)})
    left2 = pa.record_batch({"fk": pa.array(fk * 1_000_003), "q": left.column(1)})
    got, table, m = device_join(backend, stream, left2, right2, 0, 0)
    assert not table.info()[2]  # hashed
    assert_batches_equal(got, oracle.hash_join(left2, right2, 0, 0))


def test_one_key_many_build_rows(backend, stream, oracle):
    # join_timing.rs shapes: one build key with 100,000 rows, 16 probes hit
    right = pa.record_batch({"k": pa.array(np.full(100_000, 7, np.int64)), "x": pa.array(np.arange(100_000))})
    left = pa.record_batch({"k": pa.array(np.array([7] * 16 + [8] * 100, np.int64)), "y": pa.array(np.arange(116))})
    got, table, m = device_join(backend, stream, left, right, 0, 0)
    assert got.num_rows == 1_600_000
    assert_batches_equal(got, oracle.hash_join(left, right, 0, 0))

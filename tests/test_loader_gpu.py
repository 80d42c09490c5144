"""Parquet to HBM: a file written in 8,192-row row groups, loaded into one
resident batch, comes back byte for byte (values and NULLs); the k/v
GROUP BY query over a 1M-row table read that way matches the oracle."""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import datagen, loader
from oxidelake_amd.backend import FormatError
from tests.test_aggregate_gpu import assert_agg_equal
from tests.util import assert_batches_equal

pytestmark = pytest.mark.gpu


def test_round_trip(backend, stream, tmp_path):
    rng = np.random.default_rng(4)
    scalars = datagen.random_batch(rng, 50_001).select(["i", "j", "f", "g"])  # with NULLs and special floats
    vecs = pa.FixedSizeListArray.from_arrays(pa.array(rng.normal(size=50_001 * 8).astype(np.float32)), 8)
    batch = pa.RecordBatch.from_arrays(scalars.columns + [vecs], names=scalars.schema.names + ["emb"])
    path = str(tmp_path / "t.parquet")
    loader.write_parquet(batch, path)
    dev = loader.load_parquet(backend, stream, path)
    assert dev.num_rows == 50_001
    assert_batches_equal(backend.download(stream, dev), batch)


def test_columns_and_refusal(backend, stream, tmp_path):
    t = pa.table({"k": pa.array([1, 2, None], pa.int64()), "s": pa.array(["x", "y", "z"]),
                  "v": pa.array([0.5, None, 2.0])})
    path = str(tmp_path / "s.parquet")
    loader.write_parquet(t, path)
    dev = loader.load_parquet(backend, stream, path)  # the string column is left out
    assert [dev.name(i) for i in range(dev.num_columns)] == ["k", "v"]
    with pytest.raises(FormatError):
        loader.load_parquet(backend, stream, path, columns=["k", "s"])


def test_groupby_table_from_parquet(backend, stream, oracle, tmp_path):
    # the 1M-row k/v table as Parquet in 8,192-row batches, read back into
    # HBM, then SELECT k, SUM(v) WHERE k >= 2 AND v < 4.0 GROUP BY k
    n = 1_000_000
    host = backend.download(stream, datagen.table(backend, stream, datagen.groupby_table_columns(seed=42), n))
    path = str(tmp_path / "c1.parquet")
    loader.write_parquet(host, path)
    dev = loader.load_parquet(backend, stream, path)
    spec = datagen.groupby_query_spec()
    got = backend.download(stream, backend.aggregate(stream, dev, spec))
    want = oracle.aggregate(host, spec, sum_mode=1)
    assert got.num_rows == 14
    assert_agg_equal(got, want, spec)

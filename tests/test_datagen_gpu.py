"""The in-HBM synthetic table generator: counter-based determinism and the
distributions the benchmark tables are specified with."""
import numpy as np
import pytest

from oxidelake_amd import datagen, plan
from oxidelake_amd.datagen import GenColumn

pytestmark = pytest.mark.gpu


def test_prefix_of_a_larger_table_is_the_smaller_table(backend, stream):
    cols = datagen.filter_table_columns(seed=11)
    small = backend.download(stream, datagen.table(backend, stream, cols, 1000))
    big = backend.download(stream, datagen.table(backend, stream, cols, 100_000))
    assert big.slice(0, 1000).equals(small)


def test_filter_table_distributions(backend, stream):
    n = 1_000_000
    t = backend.download(stream, datagen.table(backend, stream, datagen.filter_table_columns(seed=3), n))
    a, b, c, d = (t.column(i) for i in range(4))
    a_, b_ = a.to_numpy(), b.to_numpy()
    assert a_.min() >= 0 and a_.max() < 2**40 and abs(a_.mean() / 2**39 - 1) < 0.01
    assert b_.min() == 0 and b_.max() == 999 and abs(b_.mean() - 499.5) < 2
    assert abs(c.null_count / n - 0.05) < 0.002
    cv = c.drop_null().to_numpy()
    assert abs(cv.mean()) < 0.01 and abs(cv.std() - 1) < 0.01
    ld = np.log(d.to_numpy())
    assert abs(ld.mean()) < 0.01 and abs(ld.std() - 1) < 0.01


def test_permutation_and_zipf(backend, stream):
    n = 200_000
    cols = [GenColumn("pk", plan.INT64, datagen.GEN_PERMUTATION, 0, 0, seed=5),
            GenColumn("fk", plan.INT64, datagen.GEN_ZIPF, 1.1, float(n), seed=5)]
    t = backend.download(stream, datagen.table(backend, stream, cols, n))
    pk = t.column(0).to_numpy()
    assert np.array_equal(np.sort(pk), np.arange(n))
    assert not np.array_equal(pk, np.arange(n))
    fk = t.column(1).to_numpy()
    assert fk.min() >= 0 and fk.max() < n
    # Zipf(1.1): P(rank 1) = 1 / H(n, 1.1)
    h = np.sum(np.arange(1, n + 1, dtype=np.float64) ** -1.1)
    p0 = (fk == 0).mean()
    assert abs(p0 - 1 / h) < 0.01, (p0, 1 / h)
    p1 = (fk == 1).mean()
    assert abs(p1 - 2**-1.1 / h) < 0.01


def test_normalized_mixture_vectors(backend, stream):
    cols = datagen.embedding_columns(seed=1, dim=64, centroids=16)
    t = backend.download(stream, datagen.table(backend, stream, cols, 5000))
    v = np.asarray(t.column(0).flatten().to_numpy()).reshape(5000, 64)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-5)

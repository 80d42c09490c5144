"""GPU parity of the operator over host batches (oxidelake_amd.operator):
the device path through the C ABI, with the oracle as the CPU path and as
the expected result.

Bit-exact for filter and join output (rows, order, values, NULLs) and for
integer aggregates; float sums of integer-valued inputs are exact in any
order, so they compare bit-exact too. Distances within 1e-5 relative of the
oracle's FP32 fold, as the kernel's own parity test allows.
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import plan
from oxidelake_amd.operator import GpuOperator, JoinBuild
from oxidelake_amd.plan import Aggregate, AggregateFunction as F, AggregateSpec, Comparison as Cmp, Lit, col
from tests.cpu_path import OracleCpu
from tests.util import assert_batches_equal

pytestmark = pytest.mark.gpu


def keyed(rows, p, seed=0):
    r = np.arange(rows)
    k = pa.array(r % 50, pa.int64(), mask=(r % 7 == 3))
    return pa.record_batch({f"{p}k": k, f"{p}x": pa.array(r + seed, pa.int64()),
                            f"{p}s": pa.array([f"{p}{i % 4}" for i in r]),
                            f"{p}f": pa.array(np.sin(r + seed), pa.float64())})


def concat(pieces):
    return pa.Table.from_batches(pieces).combine_chunks().to_batches()[0]


def test_operator_join_in_pieces_matches_the_oracle(backend, oracle):
    op = GpuOperator(backend, OracleCpu(oracle))
    build = JoinBuild(keyed(300, "r", 7), 0)
    parent = keyed(3_000, "l")
    proj = [1, 2, 3, 5, 6, 7]  # lx ls lf rx rs rf: the Utf8 columns gathered by row id
    for offset, size in [(0, 3_000), (1_001, 10), (3, 1)]:
        probe = parent.slice(offset, size)
        pieces = op.join_all(probe, build, 0, proj, 61)
        assert all(p.num_rows <= 61 for p in pieces)
        want = oracle.hash_join(probe, build.batch, 0, 0, proj)
        assert_batches_equal(concat(pieces) if want.num_rows else pieces[0], want)
    assert op.stats.fallback_batches == 0
    assert op.stats.rows_in == 3_000 + 10 + 1 + 300


def test_operator_filter_and_vector_distance(backend, oracle):
    op = GpuOperator(backend, OracleCpu(oracle))
    rng = np.random.default_rng(3)
    n, dim = 4_000, 16
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    vcol = pa.FixedSizeListArray.from_arrays(pa.array(vecs.ravel()), dim)
    b = pa.record_batch({"id": pa.array(np.arange(n), pa.int64()), "tag": pa.array([f"t{i}" for i in range(n)]),
                         "x": pa.array(rng.normal(size=n)), "v": vcol})
    pred = plan.cmp(col(2), Cmp.Gt, Lit.float64(0.25))
    got = op.filter_project(b, pred, [1, 0, 3])
    assert_batches_equal(got, oracle.filter_project(b, pred, [1, 0, 3]))
    q = rng.normal(size=dim).astype(np.float32)
    out = op.vector_distance(b, 3, q, plan.DistanceMetric.L2, "d")
    want = oracle.vector_distance(b, 3, q, plan.DistanceMetric.L2, "d")
    assert out.schema.names == want.schema.names
    g, w = out.column(4).to_numpy(), want.column(4).to_numpy()
    assert np.all(np.abs(g - w) <= 1e-5 * np.maximum(1.0, np.abs(w)))
    assert op.stats.fallback_batches == 0


def test_operator_aggregation_over_many_batches(backend, oracle):
    op = GpuOperator(backend, OracleCpu(oracle))
    rng = np.random.default_rng(9)
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(2)), Aggregate(F.Count, col(2)), Aggregate(F.Min, col(3)),
                             Aggregate(F.CountStar)])
    parts = []
    for i in range(5):
        m = 10_000 + 997 * i
        k = rng.integers(-20, 20, m)
        parts.append(pa.record_batch({
            "k": pa.array(k, pa.int64(), mask=rng.random(m) < 0.05),
            "s": pa.array(["x"] * m),
            "v": pa.array(rng.integers(-1000, 1000, m).astype(np.float64), mask=rng.random(m) < 0.1),
            "w": pa.array(rng.integers(-10**12, 10**12, m), pa.int64())}))
    groups = op.new_groups(parts[0], spec)
    for p in parts:
        op.aggregate_into(p, spec, groups)
    out = groups.finish()
    want = oracle.aggregate(concat(parts), spec)
    assert_batches_equal(out, want, names=False)
    st = op.stats
    assert (st.batches, st.fallback_batches, st.rows_out) == (5, 0, want.num_rows)

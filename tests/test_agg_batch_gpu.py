"""GPU parity: batched scalar aggregations (oxr_agg_batch_*) vs the CPU oracle.

Many requests of one prepared shape (the same filter and aggregates over one
global group, only literal values differing) run as one pass over a
resident table, the chain of a run replayed from a captured hipGraph. Each
request's result must be the oracle's global aggregate for its own spec
(kernels.rs:478-818): COUNT and integer SUM/MIN/MAX and float MIN/MAX bit
for bit, FP64 SUM within 1e-12 relative of the oracle's exact sum, COUNT 0
and the rest NULL over no selected row. A request's additions run in a fixed
order, so its result must also be the same bits alone, at any position of
any batch, replayed or launched directly, and from run to run.
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import autotune, datagen, plan
from oxidelake_amd.backend import ExecutionError, PlanError, Unsupported
from oxidelake_amd.plan import Aggregate, AggregateSpec, Lit, col
from oxidelake_amd.plan import AggregateFunction as F
from oxidelake_amd.plan import Comparison as Cmp
from tests.test_aggregate_gpu import assert_agg_equal

pytestmark = pytest.mark.gpu


# random_batch's columns without the vector: i Int64, j Int32, f Float32, g Float64
def table(n, seed=3):
    return datagen.random_batch(np.random.default_rng(seed), n).select([0, 1, 2, 3])


def small_spec(t, x, m, r):
    """COUNT(*), SUM(i), SUM(g*f) WHERE j < t AND (g > x OR i % m = r):
    three aggregates, 16 requests a run."""
    return AggregateSpec(-1, [Aggregate(F.CountStar, None), Aggregate(F.Sum, col(0)),
                              Aggregate(F.Sum, plan.mul(col(3), col(2)))],
                         filter=plan.and_(plan.cmp(col(1), Cmp.Lt, Lit.int32(t)),
                                          plan.or_(plan.cmp(col(3), Cmp.Gt, Lit.float64(x)),
                                                   plan.cmp(plan.mod(col(0), Lit.int64(m)), Cmp.Eq,
                                                            Lit.int64(r)))))


def wide_spec(t, y):
    """Seven aggregates (four requests a run): COUNT, MIN/MAX of every type."""
    return AggregateSpec(-1, [Aggregate(F.Count, col(2)), Aggregate(F.Min, col(2)), Aggregate(F.Max, col(3)),
                              Aggregate(F.Min, col(1)), Aggregate(F.Max, plan.add(col(0), Lit.int64(y))),
                              Aggregate(F.Sum, col(1)), Aggregate(F.Sum, col(2))],
                         filter=plan.cmp(col(1), Cmp.GtEq, Lit.int32(t)))


def small_specs(rng, n):
    return [small_spec(int(rng.integers(-25, 25)), float(rng.normal()), int(rng.integers(2, 9)),
                       int(rng.integers(0, 3))) for _ in range(n)]


def results_bits(res):
    """Every value's bits and every NULL of a run's results, for bit-equality."""
    out = []
    for rb in res:
        if isinstance(rb, Exception):
            out.append(("error",))
            continue
        row = []
        for c in rb.columns:
            vals = bytes(c.buffers()[1])[c.offset * (c.type.bit_width // 8):][:c.type.bit_width // 8]
            row.append((str(c.type), c.null_count, vals if c.null_count == 0 else None))
        out.append(tuple(row))
    return out


@pytest.fixture()
def resident(backend, stream):
    t = table(300_000)
    return t, backend.upload(stream, t)


def test_parity_against_the_oracle(backend, stream, oracle, resident):
    t, dev = resident
    rng = np.random.default_rng(11)
    b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
    assert b.max_requests == 16
    for n in (16, 5, 1):
        specs = small_specs(rng, n)
        got = b.run(stream, specs)
        assert len(got) == n
        for sp, g in zip(specs, got):
            assert_agg_equal(g, oracle.aggregate(t, sp, sum_mode=1), sp)


def test_wide_shape_parity(backend, stream, oracle, resident):
    t, dev = resident
    b = backend.prepare_agg_batch(stream, dev, wide_spec(0, 1))
    assert b.max_requests == 4
    specs = [wide_spec(-30, 5), wide_spec(0, -7), wide_spec(19, 1 << 40), wide_spec(100, 0)]
    for sp, g in zip(specs, b.run(stream, specs)):
        assert_agg_equal(g, oracle.aggregate(t, sp, sum_mode=1), sp)
    # no selected row (j >= 100): one row, COUNT 0, the rest NULL
    last = b.run(stream, [specs[3]])[0]
    assert last.num_rows == 1 and last.column(0).to_pylist() == [0]
    assert all(last.column(i).to_pylist() == [None] for i in range(1, 7))


def test_same_bits_alone_and_in_any_batch(backend, stream, resident):
    _, dev = resident
    rng = np.random.default_rng(5)
    b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
    specs = small_specs(rng, 16)
    together = results_bits(b.run(stream, specs))
    alone = [results_bits(b.run(stream, [sp]))[0] for sp in specs]
    assert together == alone
    # reversed order, and in batches of 4 and 8 (other kernels)
    assert results_bits(b.run(stream, specs[::-1])) == together[::-1]
    assert results_bits(b.run(stream, specs[:4])) == together[:4]
    assert results_bits(b.run(stream, specs[3:11])) == together[3:11]


def test_graph_replays_match_direct_launches(backend, stream, oracle, resident):
    """Replays of the captured chain, with other device work and waits between
    them, against the same chain launched directly (knob agg.batch_graph 0)."""
    t, dev = resident
    rng = np.random.default_rng(9)
    b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
    runs = [small_specs(rng, int(rng.integers(1, 17))) for _ in range(6)]
    replayed = []
    other = backend.create_stream()
    for specs in runs:
        replayed.append(results_bits(b.run(stream, specs)))
        # unrelated work on the device between replays
        backend.download(other, backend.filter_project(other, dev, plan.cmp(col(1), Cmp.Lt, Lit.int32(3)),
                                                       [col(0), col(3)]))
        other.synchronize()
    autotune.set_knob("agg.batch_graph", 0)
    try:
        direct = [results_bits(b.run(stream, specs)) for specs in runs]
    finally:
        autotune.set_knob("agg.batch_graph", -1)
    assert replayed == direct
    for sp, g in zip(runs[-1], b.run(stream, runs[-1])):
        assert_agg_equal(g, oracle.aggregate(t, sp, sum_mode=1), sp)


def test_a_dividing_request_fails_alone(backend, stream, oracle, resident):
    t, dev = resident
    b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
    specs = [small_spec(5, 0.5, 7, 1), small_spec(5, 0.5, 0, 1), small_spec(-3, 1.0, 3, 0)]
    with pytest.raises(ZeroDivisionError):
        oracle.aggregate(t, specs[1], sum_mode=1)
    got = b.run(stream, specs)
    assert isinstance(got[1], ExecutionError)
    for i in (0, 2):
        assert_agg_equal(got[i], oracle.aggregate(t, specs[i], sum_mode=1), specs[i])


def test_refusals(backend, stream, resident):
    _, dev = resident
    b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
    with pytest.raises(PlanError):
        b.run(stream, small_specs(np.random.default_rng(1), 17))
    with pytest.raises(PlanError):  # another shape: one more aggregate
        b.run(stream, [AggregateSpec(-1, small_spec(0, 0.0, 7, 3).aggregates + [Aggregate(F.CountStar, None)],
                                     filter=small_spec(0, 0.0, 7, 3).filter)])
    with pytest.raises(PlanError):  # another shape: a literal of another type
        b.run(stream, [AggregateSpec(-1, small_spec(0, 0.0, 7, 3).aggregates,
                                     filter=plan.cmp(col(1), Cmp.Lt, Lit.int64(3)))])
    assert b.run(stream, []) == []
    with pytest.raises(Unsupported):  # a group key
        backend.prepare_agg_batch(stream, dev, AggregateSpec(0, [Aggregate(F.CountStar, None)]))


def test_empty_and_ragged_tables(backend, stream, oracle):
    for n in (0, 1, 255, 1025, 70_001):
        t = table(n, seed=n)
        dev = backend.upload(stream, t)
        b = backend.prepare_agg_batch(stream, dev, small_spec(0, 0.0, 7, 3))
        specs = small_specs(np.random.default_rng(n), 3)
        for sp, g in zip(specs, b.run(stream, specs)):
            assert_agg_equal(g, oracle.aggregate(t, sp, sum_mode=1), sp)


def test_serving_filter_agg_shape(backend, stream, oracle):
    """The serving loop's filter + aggregate requests over the filter table's
    columns (a Int64, b Int32, c Float32 with NULLs, d Float64)."""
    from oxidelake_amd import serving

    cols = datagen.filter_table_columns()
    dev = datagen.table(backend, stream, cols, 400_000)
    host = backend.download(stream, dev)
    b = backend.prepare_agg_batch(stream, dev, serving.filter_agg_spec(1))
    ts = [1, 7, 50, 123, 250, 499, 500, 1000]
    for sp, g in zip([serving.filter_agg_spec(t) for t in ts],
                     b.run(stream, [serving.filter_agg_spec(t) for t in ts])):
        assert_agg_equal(g, oracle.aggregate(host, sp, sum_mode=1), sp)

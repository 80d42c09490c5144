"""The operator's host logic on a fake device (CPU only).

Mirrors the reference's operator tests (crates/oxidelake-compute/src/
operator.rs:1329-1876): the sub-batch cut and row ids, remapping, which
columns travel, fallback counters, the build side prepared once per
operator, join output in pieces of a bounded size, continuation on the CPU
path from the same match, transfer bytes of a slice, and an aggregation
handed to the device one batch at a time. The fake runs the oracle and
the expected results come from the oracle over the whole input.
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import plan
from oxidelake_amd.backend import PlanError, Unsupported
from oxidelake_amd.operator import ROW_ID, GpuOperator, JoinBuild, JoinPlan, SidePlan, transfer_bytes, travels
from oxidelake_amd.plan import Aggregate, AggregateFunction as F, AggregateSpec, Comparison as Cmp, Lit, col
from tests.cpu_path import FakeBackend, OracleCpu
from tests.util import assert_batches_equal


def batch():
    # operator.rs:1329-1348
    return pa.record_batch({
        "k": pa.array([1, None, 3, 4], pa.int64()),
        "s": pa.array(["a", "b", "c", "d"]),
        "v": pa.array([0.5, 1.5, None, 3.5], pa.float64()),
    })


def keyed(rows, p):
    # operator.rs:1543-1554: k NULL every 7th row from 3, else r % 50
    r = np.arange(rows)
    k = pa.array(r % 50, pa.int64(), mask=(r % 7 == 3))
    return pa.record_batch({f"{p}k": k, f"{p}x": pa.array(r, pa.int64()),
                            f"{p}s": pa.array([f"{p}{i % 4}" for i in r])})


def concat(pieces):
    return pa.Table.from_batches(pieces).combine_chunks().to_batches()[0] if sum(
        p.num_rows for p in pieces) else pieces[0]


def test_side_plan_cuts_eligible_columns_and_appends_row_ids():
    sp = SidePlan([2, 0], row_ids=True)
    sub = sp.sub_batch(batch())
    assert sub.schema.names == ["v", "k", ROW_ID]
    assert sp.position(0) == 1 and sp.row_id_position() == 2
    assert sub.column(2).to_pylist() == [0, 1, 2, 3]


def test_predicates_are_remapped_to_sub_batch_positions():
    sp = SidePlan([2, 0])
    p = plan.and_(plan.cmp(col(0), Cmp.Gt, Lit.int64(1)), plan.cmp(col(2), Cmp.Lt, Lit.float64(3.0)))
    assert plan.columns_of(sp.remap(p)) == [1, 0]
    with pytest.raises(PlanError):
        sp.remap(plan.cmp(col(1), Cmp.Eq, Lit.int64(0)))


def _vectors(second):
    child = pa.array([1.0, second, None, None, 5.0, 6.0], pa.float32())
    vec = pa.FixedSizeListArray.from_arrays(child, 2, mask=pa.array([False, True, False]))
    return pa.record_batch({"k": pa.array([1, 2, 3], pa.int64()), "vec": vec})


def test_a_vector_column_with_null_elements_in_valid_rows_stays_on_the_host():
    # operator.rs:1419-1466
    whole, holed = _vectors(2.0), _vectors(None)
    assert travels(whole.column(1)) and not travels(holed.column(1))
    # a NULL element only under a NULL row travels
    assert travels(_vectors(2.0).slice(1, 2).column(1))

    def sides(probe, build):
        jp = JoinPlan(probe, JoinBuild(build, 0), 0, None)
        return jp.left, jp.right

    up, gathered = SidePlan([0, 1]), SidePlan([0], row_ids=True)
    assert sides(whole, whole) == (up, up)
    assert sides(holed, whole) == (gathered, up)
    assert sides(whole, holed) == (up, gathered)


def test_filter_keeps_host_columns_on_the_host(oracle):
    fake = FakeBackend(oracle)
    op = GpuOperator(fake, OracleCpu(oracle))
    b = batch()
    out = op.filter_project(b, plan.cmp(col(0), Cmp.Gt, Lit.int64(1)), [1, 2, 0])
    assert out.schema.names == ["s", "v", "k"]
    assert out.to_pydict() == {"s": ["c", "d"], "v": [None, 3.5], "k": [3, 4]}
    assert (op.stats.batches, op.stats.fallback_batches) == (1, 0)
    # a vector column with a NULL element in a valid row is gathered by row id
    h = _vectors(None)
    out = op.filter_project(h, plan.cmp(col(0), Cmp.GtEq, Lit.int64(1)), [1, 0])
    assert out.column(0).to_pylist() == h.column(1).to_pylist() and op.stats.fallback_batches == 0


def test_a_supported_predicate_counts_no_fallback_and_a_declined_one_does(oracle):
    fp = plan.cmp(col(2), Cmp.Lt, Lit.float64(3.0))
    for floats, fallbacks in [(True, 0), (False, 1)]:
        op = GpuOperator(FakeBackend(oracle, float_predicates=floats), OracleCpu(oracle))
        out = op.filter_project(batch(), fp, [0, 2])
        assert (op.stats.batches, op.stats.fallback_batches) == (1, fallbacks)
        assert_batches_equal(out, oracle.filter_project(batch(), fp, [0, 2]))


def test_a_declined_batch_without_a_cpu_path_raises(oracle):
    op = GpuOperator(FakeBackend(oracle, float_predicates=False))
    with pytest.raises(Unsupported):
        op.filter_project(batch(), plan.cmp(col(2), Cmp.Lt, Lit.float64(3.0)), [0])
    assert op.stats.fallback_batches == 1


def test_a_sliced_batch_counts_only_the_bytes_it_moves(oracle):
    # operator.rs:1781-1805
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    sl = keyed(200_000, "t").slice(1_000, 10)
    out = op.filter_project(sl, plan.cmp(col(1), Cmp.Gt, Lit.int64(0)), [0, 1])
    assert out.num_rows == 10
    assert op.stats.bytes_h2d == 10 * 8 + 2 + 10 * 8
    assert op.stats.bytes_d2h == 10 * 8 + 2 + 10 * 8


def join_ref(oracle, probe, build, projection=None):
    return oracle.hash_join(probe, build.batch, 0, build.key, projection)


def test_join_counters_include_the_build_side_and_only_the_slices_bytes(oracle):
    # operator.rs:1574-1622
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    parent = keyed(2_000, "l")
    out_rows = 0
    for offset in (1_000, 1_010):
        out = op.join_all(parent.slice(offset, 10), build, 0, [1, 2, 4, 5], 1_000)
        assert len(out) == 1 and out[0].num_rows > 0
        assert_batches_equal(out[0], join_ref(oracle, parent.slice(offset, 10), build, [1, 2, 4, 5]))
        out_rows += out[0].num_rows
    st = op.stats
    assert st.fallback_batches == 0
    assert st.rows_in == 2 * 10 + 200
    build_bytes = 200 * 8 + 25 + 200 * 8 + 200 * 8
    probe_bytes = 10 * 8 + 2 + 10 * 8 + 10 * 8
    assert st.bytes_h2d == build_bytes + 2 * probe_bytes
    assert st.bytes_d2h == 6 * 8 * out_rows


def test_a_skewed_probe_batch_leaves_in_pieces_of_the_batch_size(oracle):
    # operator.rs:1629-1658, on the CPU path (a backend without hash joins)
    op = GpuOperator(FakeBackend(oracle, hash_join=False), OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    probe = keyed(2_000, "l")
    ref = join_ref(oracle, probe, build)
    total = ref.num_rows
    assert total > 50 * 61
    pieces = op.join_all(probe, build, 0, None, 61)
    assert len(pieces) == -(-total // 61) and max(p.num_rows for p in pieces) == 61
    assert_batches_equal(concat(pieces), ref)
    nothing = op.join_all(probe.slice(3, 1), build, 0, None, 61)  # row 3's key is NULL
    assert len(nothing) == 1 and nothing[0].num_rows == 0
    st = op.stats
    assert st.batches == st.fallback_batches == len(pieces) + 1
    assert st.rows_in == 2_000 + 200 + 1 and st.rows_out == total


def test_a_device_gathers_a_skewed_probe_batch_in_pieces(oracle):
    # operator.rs:1677-1701
    fake = FakeBackend(oracle)
    op = GpuOperator(fake, OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    probe = keyed(2_000, "l")
    ref = join_ref(oracle, probe, build)
    pieces = op.join_all(probe, build, 0, None, 61)
    assert len(pieces) == -(-ref.num_rows // 61) and all(p.num_rows <= 61 for p in pieces)
    assert_batches_equal(concat(pieces), ref)
    assert (fake.prepared, fake.probes, fake.joins) == (1, 1, len(pieces))
    assert op.stats.fallback_batches == 0 and op.stats.batches == len(pieces)


def test_a_join_prepares_its_build_side_once_per_operator(oracle):
    # operator.rs:1704-1727
    fake = FakeBackend(oracle)
    op = GpuOperator(fake, OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    parent = keyed(2_000, "l")
    for offset in (0, 500, 1_000):
        probe = parent.slice(offset, 100)
        dev = op.join_all(probe, build, 0, None, 65_536)
        ref = join_ref(oracle, probe, build)
        assert ref.num_rows > 0 and len(dev) == 1
        assert_batches_equal(dev[0], ref)
    assert (fake.prepared, fake.probes, fake.uploads) == (1, 3, 1 + 3)


def test_a_backend_that_cannot_prepare_a_join_is_asked_once(oracle):
    # operator.rs:1730-1752
    fake = FakeBackend(oracle, declines_prepare=True)
    op = GpuOperator(fake, OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    probe = keyed(100, "l")
    ref = join_ref(oracle, probe, build)
    for _ in range(3):
        out = op.join_all(probe, build, 0, None, 65_536)
        assert len(out) == 1
        assert_batches_equal(out[0], ref)
    assert op.stats.fallback_batches == 3
    assert (fake.uploads, fake.probes) == (1, 0)


def test_a_probe_batch_finishes_on_the_cpu_when_the_device_declines_a_piece(oracle):
    # operator.rs:1755-1778
    fake = FakeBackend(oracle, gathers_before_declining=2)
    op = GpuOperator(fake, OracleCpu(oracle))
    build = JoinBuild(keyed(200, "r"), 0)
    probe = keyed(2_000, "l")
    ref = join_ref(oracle, probe, build, [1, 2, 4, 5])
    pieces = op.join_all(probe, build, 0, [1, 2, 4, 5], 61)
    assert all(p.num_rows <= 61 for p in pieces)
    assert_batches_equal(concat(pieces), ref)
    assert op.stats.fallback_batches == len(pieces) - 2


def test_join_projection_out_of_range_is_a_plan_error(oracle):
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    with pytest.raises(PlanError):
        op.hash_join(keyed(10, "l"), JoinBuild(keyed(10, "r"), 0), 0, [7])


SPEC = AggregateSpec(0, [Aggregate(F.Sum, col(2)), Aggregate(F.Count, col(2)), Aggregate(F.Max, col(0))])


def test_the_aggregation_streams_batches_to_the_device(oracle):
    # operator.rs:1807-1846: three 4-row batches are three 4-row updates,
    # equal to the reference over all twelve rows; rows out are the groups
    fake = FakeBackend(oracle)
    op = GpuOperator(fake, OracleCpu(oracle))
    groups = op.new_groups(batch(), SPEC)
    for _ in range(3):
        op.aggregate_into(batch(), SPEC, groups)
    out = groups.finish()
    assert fake.aggregate_rows == [4, 4, 4]
    all_rows = pa.Table.from_batches([batch()] * 3).combine_chunks().to_batches()[0]
    assert_batches_equal(out, oracle.aggregate(all_rows, SPEC), names=False)
    st = op.stats
    assert (st.batches, st.rows_in, st.rows_out, st.fallback_batches) == (3, 12, 4, 0)
    # only the key and the aggregate inputs travel: k and v, not s
    assert st.bytes_h2d == 3 * (4 * 8 + 1 + 4 * 8 + 1)


def test_a_declined_aggregate_batch_is_merged_with_the_device_groups(oracle):
    fake = FakeBackend(oracle, declines_update={1})
    op = GpuOperator(fake, OracleCpu(oracle))
    rng = np.random.default_rng(5)
    parts = [pa.record_batch({"k": pa.array(rng.integers(0, 6, 50), pa.int64()), "s": pa.array(["x"] * 50),
                              "v": pa.array(rng.integers(-9, 9, 50).astype(np.float64))}) for _ in range(3)]
    groups = op.new_groups(parts[0], SPEC)
    for p in parts:
        op.aggregate_into(p, SPEC, groups)
    out = groups.finish()
    assert op.stats.fallback_batches == 1
    want = oracle.aggregate(pa.Table.from_batches(parts).combine_chunks().to_batches()[0], SPEC)
    assert_batches_equal(out, want, names=False)


def test_every_operator_counts_its_fallbacks(oracle):
    # operator.rs:1849-1875
    op = GpuOperator(FakeBackend(oracle, aggregate=False), OracleCpu(oracle))
    op.aggregate(batch(), AggregateSpec(0, [Aggregate(F.Count, col(2))]))
    assert op.stats.fallback_batches == 1
    # a vector column the device cannot take runs the CPU path
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    h = _vectors(None)
    out = op.vector_distance(h, 1, [1.0, 0.0], plan.DistanceMetric.L2, "d")
    assert op.stats.fallback_batches == 1
    assert out.schema.names == ["k", "vec", "d"]


def test_an_empty_aggregation_finishes(oracle):
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    out = op.new_groups(batch(), SPEC).finish()
    assert out.num_rows == 0 and out.num_columns == 4


def test_vector_distance_appends_one_column(oracle):
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    w = _vectors(2.0)
    out = op.vector_distance(w, 1, [1.0, 0.0], plan.DistanceMetric.L2, "d")
    assert out.schema.names == ["k", "vec", "d"] and out.schema.field(2).type == pa.float32()
    assert_batches_equal(out, oracle.vector_distance(w, 1, [1.0, 0.0], plan.DistanceMetric.L2, "d"))
    assert op.stats.bytes_h2d == transfer_bytes(w.select(["vec"]))


def test_no_predicate_empty_batches_and_empty_probes(oracle):
    op = GpuOperator(FakeBackend(oracle), OracleCpu(oracle))
    b = batch()
    # no predicate keeps every row; the string column comes back by row id
    out = op.filter_project(b, None, [1, 0])
    assert out.to_pydict() == {"s": ["a", "b", "c", "d"], "k": [1, None, 3, 4]}
    # an empty batch, through every operator
    e = b.slice(0, 0)
    assert op.filter_project(e, plan.cmp(col(0), Cmp.Gt, Lit.int64(1)), [1, 0]).num_rows == 0
    build = JoinBuild(keyed(20, "r"), 0)
    pieces = op.join_all(keyed(10, "l").slice(0, 0), build, 0, None, 7)
    assert len(pieces) == 1 and pieces[0].num_rows == 0 and pieces[0].num_columns == 6
    groups = op.new_groups(b, SPEC)
    op.aggregate_into(e, SPEC, groups)
    op.aggregate_into(b, SPEC, groups)
    assert_batches_equal(groups.finish(), oracle.aggregate(b, SPEC), names=False)
    assert op.stats.fallback_batches == 0

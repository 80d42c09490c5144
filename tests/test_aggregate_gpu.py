"""GPU parity: grouped aggregation vs the CPU oracle.

Keys, COUNT, integer SUM/MIN/MAX and float MIN/MAX: bit-exact. Float SUM:
within 1e-12 relative of the oracle's compensated (exact) sum — the device
adds in a different order than the CPU's sequential fold, and that order is
the only divergence (crates/oxidelake-device/src/cpu/kernels.rs:478-818).
"""
import math
import struct

import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import datagen, plan
from oxidelake_amd.backend import AggState
from oxidelake_amd.plan import AggregateFunction as F
from oxidelake_amd.plan import Aggregate, AggregateSpec, Lit, col
from oxidelake_amd.plan import Comparison as Cmp
from tests.test_oracle_golden import EDGE_ROWS, EDGE_SPEC, edge_batch
from tests.util import assert_arrays_equal

pytestmark = pytest.mark.gpu

SUM_REL = 1e-12


SPILL_KNOBS = {"agg.range_rows": 0, "agg.spill_rows": 0, "agg.dense": 0, "agg.spill_pct": 100}
# hot: the spill on, and a sampled hot-group map from any input size
HOT_KNOBS = {**SPILL_KNOBS, "agg.hot_rows": 1}


@pytest.fixture(autouse=True, params=["generated", "prepared", "exact", "dense", "spill", "hot", "interpreter"])
def path(request):
    # every test runs on the kernel generated per aggregation shape (with
    # literals as constants; prepared: literals as kernel arguments; exact:
    # compensated FP64 sums; dense: every key range measured, so small ones
    # take the per-lane LDS kernel; spill: rows the block table does not
    # hold are spilled, partitioned and folded in LDS) and on the
    # interpreter kernel
    from oxidelake_amd import autotune

    autotune.set_knob("agg.jit", 0 if request.param == "interpreter" else 1)
    autotune.set_knob("jit.params", 1 if request.param == "prepared" else 0)
    autotune.set_knob("agg.exact_f64", 1 if request.param == "exact" else 0)
    autotune.set_knob("agg.range_rows", 0 if request.param == "dense" else -1)
    if request.param in ("spill", "hot"):
        for k, v in (SPILL_KNOBS if request.param == "spill" else HOT_KNOBS).items():
            autotune.set_knob(k, v)
    yield request.param
    for k in ("agg.jit", "jit.params", "agg.exact_f64", "agg.range_rows", *HOT_KNOBS):
        autotune.set_knob(k, -1)


def fbits(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


def assert_agg_equal(got: pa.RecordBatch, want: pa.RecordBatch, spec: AggregateSpec):
    assert got.num_rows == want.num_rows, f"{got.num_rows} groups != {want.num_rows}"
    assert got.num_columns == want.num_columns
    key_cols = 0 if spec.group_by < 0 else 1
    if key_cols:
        assert_arrays_equal(got.column(0), want.column(0), "key")
    for i, a in enumerate(spec.aggregates):
        g, w = got.column(key_cols + i), want.column(key_cols + i)
        assert g.type == w.type, (i, g.type, w.type)
        if a.func == F.Sum and w.type == pa.float64():
            gv, wv = g.to_pylist(), w.to_pylist()
            for r, (x, y) in enumerate(zip(gv, wv)):
                assert (x is None) == (y is None), (i, r, x, y)
                if x is None:
                    continue
                if math.isnan(y) or math.isinf(y):
                    assert (math.isnan(x) and math.isnan(y)) or x == y, (i, r, x, y)
                else:
                    assert abs(x - y) <= SUM_REL * max(abs(y), 1e-300), (i, r, x, y, abs(x - y) / max(abs(y), 1e-300))
        else:
            assert_arrays_equal(g, w, f"aggregate {i} ({got.schema.field(key_cols + i).name})")


def device_aggregate(backend, stream, batch, spec):
    dev = backend.upload(stream, batch)
    return backend.download(stream, backend.aggregate(stream, dev, spec))


def test_reference_known_answer(backend, stream, oracle):
    # kernels.rs:1123-1158
    b = pa.record_batch({"k": pa.array([1, 2, None, 2, 5], pa.int64()),
                         "v": pa.array([0.5, None, 2.5, 3.5, -1.0], pa.float64())})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)),
                             Aggregate(F.Min, col(1)), Aggregate(F.Max, col(0))])
    got = device_aggregate(backend, stream, b, spec)
    assert got.column(0).to_pylist() == [1, 2, 5, None]
    assert got.column(1).to_pylist() == [0.5, 3.5, -1.0, 2.5]
    assert got.column(2).to_pylist() == [1, 1, 1, 1]
    assert got.column(4).to_pylist() == [1, 2, 5, None]
    assert got.schema.field(1).name == "SUM(v)" and got.schema.field(4).name == "MAX(k)"
    assert_agg_equal(got, oracle.aggregate(b, spec, sum_mode=1), spec)


def test_edge_values_wrap_and_total_order(backend, stream, oracle):
    # kernels.rs:1209-1288: wrapped sums, NaN of both signs, signed zeros
    b = edge_batch(EDGE_ROWS)
    got = device_aggregate(backend, stream, b, EDGE_SPEC)
    assert got.column(1).to_pylist() == [-7, 4, None, -(2**63) + 7, -2]
    assert_agg_equal(got, oracle.aggregate(b, EDGE_SPEC, sum_mode=1), EDGE_SPEC)


def test_sum_overflow_rows(backend, stream):
    # conformance.rs:306-356
    I64MAX = 2**63 - 1
    rows = [(1, I64MAX - 1), (2, 5), (1, I64MAX - 1), (3, -(2**63)), (4, I64MAX), (3, -1),
            (None, I64MAX), (4, 1), (None, I64MAX), (4, -2)]
    b = pa.record_batch({"k": pa.array([r[0] for r in rows], pa.int64()),
                         "x": pa.array([r[1] for r in rows], pa.int64())})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1))])
    got = device_aggregate(backend, stream, b, spec)
    rendered = ["|".join("NULL" if v is None else str(v) for v in row)
                for row in zip(*[c.to_pylist() for c in got.columns])]
    assert rendered == ["1|-4|2", "2|5|1", f"3|{I64MAX}|2", f"4|{I64MAX - 1}|3", "NULL|-2|2"]


def test_empty_batch_has_no_groups(backend, stream):
    b = pa.record_batch({"k": pa.array([], pa.int64()), "v": pa.array([], pa.float64())})
    got = device_aggregate(backend, stream, b, AggregateSpec(0, [Aggregate(F.Count, col(1))]))
    assert got.num_rows == 0 and got.num_columns == 2


def test_int64_min_key_is_a_group(backend, stream, oracle):
    b = pa.record_batch({"k": pa.array([-(2**63), 5, -(2**63), None, 5], pa.int64()),
                         "x": pa.array([1, 2, 3, 4, 5], pa.int64())})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.CountStar)])
    got = device_aggregate(backend, stream, b, spec)
    assert got.column(0).to_pylist() == [-(2**63), 5, None]
    assert got.column(1).to_pylist() == [4, 7, 4]
    assert_agg_equal(got, oracle.aggregate(b, spec), spec)


SPECS = [
    AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)), Aggregate(F.Min, col(1)),
                      Aggregate(F.Max, col(1))]),
    AggregateSpec(1, [Aggregate(F.Sum, col(3)), Aggregate(F.Min, col(3)), Aggregate(F.Max, col(3)),
                      Aggregate(F.CountStar)]),
    AggregateSpec(0, [Aggregate(F.Sum, col(2)), Aggregate(F.Min, col(2)), Aggregate(F.Max, col(2)),
                      Aggregate(F.Count, col(2))]),
    AggregateSpec(1, [Aggregate(F.Sum, plan.mul(col(0), col(1))), Aggregate(F.Sum, plan.mul(col(2), col(3))),
                      Aggregate(F.CountStar), Aggregate(F.Min, plan.add(col(1), Lit.int32(1))),
                      Aggregate(F.Max, col(0)), Aggregate(F.Count, col(3))]),
    AggregateSpec(-1, [Aggregate(F.Sum, col(0)), Aggregate(F.Count, col(3)), Aggregate(F.Max, col(3))]),
    AggregateSpec(0, [Aggregate(F.Sum, col(3))],
                  filter=plan.and_(plan.cmp(col(1), Cmp.GtEq, Lit.int64(-5)),
                                   plan.cmp(col(3), Cmp.Lt, Lit.float64(4.0)))),
]


@pytest.mark.parametrize("n", [1, 7, 64, 1000, 50_000])
@pytest.mark.parametrize("si", range(len(SPECS)))
def test_random_parity(backend, stream, oracle, n, si):
    rng = np.random.default_rng(100 + n)
    batch = datagen.random_batch(rng, n).select([0, 1, 2, 3])
    spec = SPECS[si]
    got = device_aggregate(backend, stream, batch, spec)
    want = oracle.aggregate(batch, spec, sum_mode=1)
    assert_agg_equal(got, want, spec)


def test_many_groups_and_streaming_state(backend, stream, oracle):
    rng = np.random.default_rng(8)
    n = 200_000
    k = rng.integers(0, 60_000, n).astype(np.int64)
    x = rng.integers(-(2**40), 2**40, n).astype(np.int64)
    v = rng.random(n) * 1000
    batch = pa.record_batch({"k": pa.array(k), "x": pa.array(x), "v": pa.array(v)})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Sum, col(2)), Aggregate(F.Min, col(1)),
                             Aggregate(F.Max, col(2)), Aggregate(F.CountStar)])
    want = oracle.aggregate(batch, spec, sum_mode=1)
    got = device_aggregate(backend, stream, batch, spec)
    assert_agg_equal(got, want, spec)
    # the same rows streamed through a resident state in uneven batches
    dev_batches = [backend.upload(stream, batch.slice(a, b - a))
                   for a, b in [(0, 7), (7, 50_000), (50_000, 50_001), (50_001, n)]]
    st = AggState(backend, stream, dev_batches[0], spec)
    for d in dev_batches:
        st.update(stream, d)
    got = backend.download(stream, st.finish(stream))
    assert_agg_equal(got, want, spec)


def test_groupby_query(backend, stream, oracle):
    # SELECT k, SUM(v) FROM t WHERE k >= 2 AND v < 4.0 GROUP BY k over 1M rows
    rng = np.random.default_rng(42)
    n = 1_000_000
    k = rng.integers(0, 16, n).astype(np.int32)
    v = rng.random(n) * 8
    batch = pa.record_batch({"k": pa.array(k), "v": pa.array(v)})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1))],
                         filter=plan.and_(plan.cmp(col(0), Cmp.GtEq, Lit.int32(2)),
                                          plan.cmp(col(1), Cmp.Lt, Lit.float64(4.0))))
    got = device_aggregate(backend, stream, batch, spec)
    want = oracle.aggregate(batch, spec, sum_mode=1)
    assert got.num_rows == 14
    assert got.column(0).type == pa.int32()
    assert_agg_equal(got, want, spec)
    exact = [math.fsum(v[(k == g) & (v < 4.0)]) for g in range(2, 16)]
    assert_agg_equal(got, pa.record_batch({"k": got.column(0), "SUM(v)": pa.array(exact)}), spec)


def test_global_aggregate_over_no_rows_is_one_row(backend, stream, oracle):
    b = pa.record_batch({"x": pa.array([], pa.int64()), "v": pa.array([], pa.float64())})
    spec = AggregateSpec(-1, [Aggregate(F.Sum, col(0)), Aggregate(F.Count, col(1)), Aggregate(F.Max, col(1)),
                              Aggregate(F.CountStar)])
    got = device_aggregate(backend, stream, b, spec)
    assert [c.to_pylist() for c in got.columns] == [[None], [0], [None], [0]]
    assert_agg_equal(got, oracle.aggregate(b, spec), spec)


# global aggregates (no key): every aggregate kind, filtered, over NULLs and
# the float specials of random_batch (i Int64, j Int32, f Float32, g Float64)
GLOBAL_SPECS = [
    AggregateSpec(-1, [Aggregate(F.CountStar), Aggregate(F.Sum, col(0)), Aggregate(F.Sum, plan.mul(col(2), col(3))),
                       Aggregate(F.Min, col(2)), Aggregate(F.Max, col(1)), Aggregate(F.Count, col(2)),
                       Aggregate(F.Max, col(3)), Aggregate(F.Min, col(0))],
                  filter=plan.cmp(col(1), Cmp.Lt, Lit.int32(3))),
    AggregateSpec(-1, [Aggregate(F.Sum, col(3)), Aggregate(F.Sum, col(1)), Aggregate(F.Min, col(3))]),
    AggregateSpec(-1, [Aggregate(F.Count, col(0)), Aggregate(F.Sum, col(2))],
                  filter=plan.cmp(col(1), Cmp.Gt, Lit.int32(100))),  # no row selected
]


@pytest.mark.parametrize("n", [0, 1, 1000, 2_000_003])
def test_global_aggregates_take_the_keyless_kernel(backend, stream, oracle, path, n):
    """A global aggregate runs the keyless kernel (per-lane accumulators,
    one table update per block) on every generated path."""
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    batch = datagen.random_batch(np.random.default_rng(n + 5), n).select([0, 1, 2, 3])
    dev = backend.upload(stream, batch)
    for spec in GLOBAL_SPECS:
        lib.oxr_profile_enable(1)
        lib.oxr_profile_reset()
        try:
            got = backend.download(stream, backend.aggregate(stream, dev, spec))
            # (no launch at all over no rows)
            assert autotune.launches("aggregate_keyless") == (0 if path == "interpreter" or n == 0 else 1)
        finally:
            lib.oxr_profile_enable(0)
            lib.oxr_profile_reset()
        assert_agg_equal(got, oracle.aggregate(batch, spec, sum_mode=1), spec)


def test_keyless_streams_and_matches_the_block_table(backend, stream, oracle, path):
    """The keyless kernel over a resident state fed in uneven batches, and
    against the block-table kernel (knob agg.keyless 0) on the same input:
    integers and float MIN/MAX bit for bit, FP64 sums within 1e-12."""
    from oxidelake_amd import autotune

    n = 400_000
    batch = datagen.random_batch(np.random.default_rng(77), n).select([0, 1, 2, 3])
    spec = GLOBAL_SPECS[0]
    want = oracle.aggregate(batch, spec, sum_mode=1)
    parts = [backend.upload(stream, batch.slice(a, b - a)) for a, b in [(0, 3), (3, 70_000), (70_000, n)]]
    st = AggState(backend, stream, parts[0], spec)
    for d in parts:
        st.update(stream, d)
    assert_agg_equal(backend.download(stream, st.finish(stream)), want, spec)
    dev = backend.upload(stream, batch)
    keyless = backend.download(stream, backend.aggregate(stream, dev, spec))
    autotune.set_knob("agg.keyless", 0)
    try:
        table = backend.download(stream, backend.aggregate(stream, dev, spec))
    finally:
        autotune.set_knob("agg.keyless", -1)
    assert_agg_equal(keyless, table, spec)


def test_keyless_division_by_zero_fails(backend, stream, oracle, path):
    from oxidelake_amd.backend import ExecutionError

    b = pa.record_batch({"x": pa.array([5, 7, 9, 11], pa.int64()), "y": pa.array([1, 2, 0, 3], pa.int64())})
    spec = AggregateSpec(-1, [Aggregate(F.Sum, plan.mod(col(0), col(1)))])
    with pytest.raises(ZeroDivisionError):
        oracle.aggregate(b, spec)
    with pytest.raises(ExecutionError):
        device_aggregate(backend, stream, b, spec)
    # the same expression over rows the filter leaves out of the zero
    ok = AggregateSpec(-1, spec.aggregates, filter=plan.cmp(col(1), Cmp.NotEq, Lit.int64(0)))
    assert_agg_equal(device_aggregate(backend, stream, b, ok), oracle.aggregate(b, ok), ok)


def test_key_range_is_measured_once_per_batch(backend, stream, oracle, path):
    """The key column's value range (table sizing, kernel choice) is kept
    with the batch: measured by the first update over a batch, again for a
    new batch, every time with knob agg.range_cache 0; the same results."""
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    autotune.set_knob("agg.range_rows", 0)  # measure at any size (reset by the fixture)
    rng = np.random.default_rng(44)
    n = 120_000
    b = pa.record_batch({"k": pa.array(rng.integers(-20, 30, n).astype(np.int64), mask=rng.random(n) < 0.05),
                         "v": pa.array(rng.normal(size=n))})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.CountStar), Aggregate(F.Max, col(1))])
    want = oracle.aggregate(b, spec, sum_mode=1)
    dev = backend.upload(stream, b)
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        outs = [backend.download(stream, backend.aggregate(stream, dev, spec)) for _ in range(3)]
        assert autotune.launches("key_range") == 1
        other = backend.upload(stream, b)
        outs.append(backend.download(stream, backend.aggregate(stream, other, spec)))
        assert autotune.launches("key_range") == 2
        autotune.set_knob("agg.range_cache", 0)
        outs.append(backend.download(stream, backend.aggregate(stream, dev, spec)))
        outs.append(backend.download(stream, backend.aggregate(stream, dev, spec)))
        assert autotune.launches("key_range") == 4
    finally:
        autotune.set_knob("agg.range_cache", -1)
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    for o in outs:
        assert_agg_equal(o, want, spec)


def test_hot_keys_fold_across_the_wave(backend, stream, oracle):
    # one key on most rows (the matching round folds it per wave), NULL keys
    # and NULL inputs mixed in, a cold tail of distinct keys
    rng = np.random.default_rng(21)
    n = 300_000
    k = np.where(rng.random(n) < 0.7, 7, rng.integers(0, 100_000, n)).astype(np.int64)
    x = rng.integers(-1000, 1000, n).astype(np.int64)
    v = rng.normal(size=n)
    b = pa.record_batch({"k": pa.array(k, mask=rng.random(n) < 0.05),
                         "x": pa.array(x, mask=rng.random(n) < 0.1), "v": pa.array(v)})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Sum, col(2)), Aggregate(F.Min, col(2)),
                             Aggregate(F.Max, col(1)), Aggregate(F.Count, col(1))])
    assert_agg_equal(device_aggregate(backend, stream, b, spec), oracle.aggregate(b, spec, sum_mode=1), spec)


@pytest.mark.parametrize("rows", [1, 4, 16, 32])
def test_every_rows_per_lane(backend, stream, oracle, rows, path):
    from oxidelake_amd import autotune

    autotune.set_knob("agg.rows", rows)
    try:
        rng = np.random.default_rng(31)
        batch = datagen.random_batch(rng, 20_001).select([0, 1, 2, 3])
        for spec in (SPECS[0], SPECS[3], SPECS[5]):
            got = device_aggregate(backend, stream, batch, spec)
            assert_agg_equal(got, oracle.aggregate(batch, spec, sum_mode=1), spec)
    finally:
        autotune.set_knob("agg.rows", -1)


def test_generated_kernel_is_the_one_that_runs(backend, stream, path):
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    dev = datagen.table(backend, stream, datagen.groupby_table_columns(seed=3), 10_000)
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        backend.aggregate(stream, dev, datagen.groupby_query_spec())
        want = {"interpreter": (0, 0, 1), "dense": (0, 1, 0)}.get(path, (1, 0, 0))
        got = (autotune.launches("aggregate"), autotune.launches("aggregate_dense"), autotune.launches("aggregate_vm"))
        assert got == want
    finally:
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()


def _ulps(x, y):
    return abs(x - y) / math.ulp(abs(y)) if y else abs(x)


@pytest.mark.parametrize("group_by", [0, -1])
def test_float_sums_are_compensated(backend, stream, group_by, path):
    # 1e16 + 1 rounds back to 1e16 in FP64, so a plain sum of
    # [1e16, 1 x 5000, -1e16] in any order that meets 1e16 first loses ones;
    # the compensated device sum is exact. Key 1 holds a cancelling series
    # (N(0,1) x lognormal): its sum stays within a few ulps of the exact
    # (math.fsum) sum, in whatever order the atomics land.
    rng = np.random.default_rng(8)
    ones = 5000
    v0 = np.concatenate([[1e16, -1e16], np.ones(ones)])
    v1 = rng.normal(size=300_000) * rng.lognormal(size=300_000)
    k = np.concatenate([np.zeros(len(v0), np.int64), np.ones(len(v1), np.int64)])
    v = np.concatenate([v0, v1])
    perm = rng.permutation(len(v))
    k, v = k[perm], v[perm]
    if group_by < 0:  # the global group over key 0's rows only
        m = k == 0
        k, v = k[m], v[m]
    batch = pa.record_batch({"k": pa.array(k), "v": pa.array(v)})
    spec = AggregateSpec(group_by, [Aggregate(F.Sum, col(1))])
    from oxidelake_amd import autotune

    autotune.set_knob("agg.exact_f64", 1)  # on every path (reset by the fixture)
    got = device_aggregate(backend, stream, batch, spec)
    sums = got.column(got.num_columns - 1).to_pylist()
    assert sums[0] == float(ones)
    if group_by == 0:
        assert got.column(0).to_pylist() == [0, 1]
        want = math.fsum(v1)
        assert _ulps(sums[1], want) <= 4, (sums[1], want)
    # the same twice: no run-to-run drift
    again = device_aggregate(backend, stream, batch, spec)
    assert again.column(again.num_columns - 1).to_pylist() == sums


@pytest.mark.parametrize("keys", ["int32", "int64_negative", "nullable", "wide"])
def test_dense_small_key_ranges(backend, stream, oracle, path, keys):
    # the per-lane LDS kernel: Int32 and negative Int64 keys, a NULL-key slot,
    # and a range too wide for it (the hashed kernel takes that one)
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    rng = np.random.default_rng(17)
    n = 150_000
    if keys == "int32":
        k = pa.array(rng.integers(5, 12, n).astype(np.int32))
    elif keys == "int64_negative":
        k = pa.array(rng.integers(-9, -1, n).astype(np.int64))
    elif keys == "nullable":
        k = pa.array(rng.integers(0, 6, n).astype(np.int64), mask=rng.random(n) < 0.1)
    else:
        k = pa.array(rng.integers(0, 40, n).astype(np.int64))
    v = pa.array(rng.normal(size=n), mask=rng.random(n) < 0.05)
    w = pa.array(rng.integers(-1000, 1000, n).astype(np.int32))
    batch = pa.record_batch({"k": k, "v": v, "w": w})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)), Aggregate(F.Min, col(2)),
                             Aggregate(F.Max, col(1))],
                         filter=plan.cmp(col(2), Cmp.Gt, Lit.int32(-900)))
    autotune.set_knob("agg.range_rows", 0)  # measure the range on every path
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        got = device_aggregate(backend, stream, batch, spec)
        dense = autotune.launches("aggregate_dense")
    finally:
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    want = oracle.aggregate(batch, spec, sum_mode=1)
    assert_agg_equal(got, want, spec)
    if path == "interpreter":
        assert dense == 0
    else:
        # 4 aggregates x (range + NULL slot) slots fit 32 unless the range is wide
        # (twice that in exact mode: too many for every case here; the spill
        # path turns the dense kernel off)
        assert dense == (0 if keys == "wide" or path in ("exact", "spill", "hot") else 1), dense


@pytest.mark.parametrize("pct", [1, 100])
def test_spill_many_groups(backend, stream, oracle, path, pct):
    # 300K distinct keys over 1M rows: most rows miss the block tables and
    # are spilled, partitioned and folded (pct 1: the spill regions fill up
    # and the rest take the HBM atomics, both in one update)
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    rng = np.random.default_rng(23)
    n = 1_000_000
    k = pa.array(rng.integers(-150_000, 150_000, n).astype(np.int64), mask=rng.random(n) < 0.01)
    # positive values: a sum of a few cancelling terms is ill-conditioned and
    # any other order of additions leaves it more than 1e-12 relative apart
    v = pa.array(rng.uniform(0.5, 5, n), mask=rng.random(n) < 0.05)
    w = pa.array(rng.integers(0, 1000, n).astype(np.int32))
    batch = pa.record_batch({"k": k, "v": v, "w": w})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.CountStar, None), Aggregate(F.Max, col(2)),
                             Aggregate(F.Sum, col(2))], filter=plan.cmp(col(2), Cmp.Lt, Lit.int32(900)))
    for kk, vv in (HOT_KNOBS if path == "hot" else SPILL_KNOBS).items():
        autotune.set_knob(kk, vv)
    autotune.set_knob("agg.spill_pct", pct)
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        got = device_aggregate(backend, stream, batch, spec)
        spills = autotune.launches("aggregate_spill")
        hot = autotune.launches("aggregate_hot")
    finally:
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    assert_agg_equal(got, oracle.aggregate(batch, spec, sum_mode=1), spec)
    assert spills == (0 if path in ("interpreter", "exact") else 1)
    assert hot == (1 if path == "hot" else 0)


def test_groups_hint_below_the_distinct_keys(backend, stream, oracle, path):
    # a hint far below the keys the batches bring only pre-sizes the table:
    # every update grows it first, so no row is lost or counted twice
    from oxidelake_amd import autotune

    autotune.set_knob("agg.range_rows", 0)  # measured key ranges on every batch size
    rng = np.random.default_rng(41)
    batches = []
    for n in (5_000, 70_000, 200_000):
        k = pa.array(rng.integers(0, 50_000, n).astype(np.int64), mask=rng.random(n) < 0.02)
        v = pa.array(rng.integers(-1000, 1000, n).astype(np.int64))
        batches.append(pa.record_batch({"k": k, "v": v}))
    spec = AggregateSpec(0, [Aggregate(F.CountStar, None), Aggregate(F.Sum, col(1)), Aggregate(F.Max, col(1))],
                         groups_hint=4)
    dev = [backend.upload(stream, b) for b in batches]
    st = backend.aggregate_state(stream, dev[0], spec)
    for d in dev:
        st.update(stream, d)
    got = backend.download(stream, st.finish(stream))
    want = oracle.aggregate(pa.Table.from_batches(batches).combine_chunks().to_batches()[0], spec, sum_mode=1)
    assert_agg_equal(got, want, spec)


@pytest.mark.parametrize("null_keys", [False, True])
def test_hot_groups_of_a_skewed_key(backend, stream, oracle, path, null_keys):
    # a Zipf-like key over 200K values: the sampled hot groups take most
    # rows in LDS, the rest spill or go to the HBM table; every group's
    # result is the same whichever way its rows went
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    rng = np.random.default_rng(57)
    n = 3_000_000
    ranks = np.minimum((rng.pareto(0.4, n) * 3).astype(np.int64), 199_999)
    perm = rng.permutation(200_000).astype(np.int64) - 50_000
    k = pa.array(perm[ranks], mask=(rng.random(n) < 0.02) if null_keys else None)
    x = pa.array(rng.integers(-500, 500, n).astype(np.int64), mask=rng.random(n) < 0.03)
    v = pa.array(rng.uniform(0.5, 5, n))
    batch = pa.record_batch({"k": k, "x": x, "v": v})
    spec = AggregateSpec(0, [Aggregate(F.CountStar, None), Aggregate(F.Sum, col(1)), Aggregate(F.Min, col(1)),
                             Aggregate(F.Sum, col(2))], filter=plan.cmp(col(1), Cmp.NotEq, Lit.int64(7)))
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        got = device_aggregate(backend, stream, batch, spec)
        hot = autotune.launches("aggregate_hot")
    finally:
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    assert_agg_equal(got, oracle.aggregate(batch, spec, sum_mode=1), spec)
    assert hot == (1 if path == "hot" else 0)

"""GPU parity: fused filter + projection vs the CPU oracle.

Bit-exact: the selected rows, their order, every projected value (bit
pattern) and every NULL must equal the oracle's restatement of
crates/oxidelake-device/src/cpu/kernels.rs:99-152. Three device paths are
checked: the kernel generated per query shape, with literals compiled in
("generated") or passed as kernel arguments ("prepared"), and the
expression interpreter kernel ("interpreter").
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import autotune, datagen, plan
from oxidelake_amd.backend import ExecutionError, Unsupported
from oxidelake_amd.plan import Comparison as Cmp
from oxidelake_amd.plan import Lit, Predicate, col
from tests.util import assert_batches_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["generated", "prepared", "persistent", "interpreter"])
def path(request):
    # generated: the generated kernel, one tile per workgroup launch;
    # prepared: the same with literals passed as arguments; persistent: the
    # persistent generated kernel (tiles by ticket, the next tile's loads in
    # flight)
    autotune.set_knob("filter.jit", 0 if request.param == "interpreter" else 1)
    autotune.set_knob("jit.params", 1 if request.param == "prepared" else 0)
    autotune.set_knob("filter.persistent", 1 if request.param == "persistent" else 0)
    yield request.param
    autotune.set_knob("filter.jit", -1)
    autotune.set_knob("jit.params", -1)
    autotune.set_knob("filter.persistent", -1)


@pytest.fixture
def small_tiles():
    # 4 rows per lane: the random grid compiles one kernel per query shape,
    # and a 16-row kernel takes twice as long to compile
    autotune.set_knob("filter.rows", 4)
    yield
    autotune.set_knob("filter.rows", -1)


def run_both(backend, stream, oracle, batch, pred, proj, names=None):
    dev = backend.upload(stream, batch)
    out = backend.filter_project(stream, dev, pred, proj, names)
    got = backend.download(stream, out)
    want = oracle.filter_project(batch, pred, proj, names)
    return got, want


def ref_batch():
    # kernels.rs:938-973
    return pa.record_batch({
        "k": pa.array([1, 2, None, 2, 5], pa.int64()),
        "v": pa.array([0.5, None, 2.5, 3.5, -1.0], pa.float64()),
    })


def test_reference_known_answer(backend, stream, oracle, path):
    # kernels.rs:976-988: output k = [2, 5]
    pred = Predicate.and_(Predicate.compare(0, Cmp.GtEq, Lit.int64(2)),
                          Predicate.compare(1, Cmp.Lt, Lit.float64(4.0)))
    got, want = run_both(backend, stream, oracle, ref_batch(), pred, [1, 0])
    assert got.column(1).to_pylist() == [2, 5]
    assert_batches_equal(got, want)


def test_float_order_as_datafusion(backend, stream, oracle, path):
    # kernels.rs:990-1021
    vals = [float("nan"), -float("nan"), float("inf"), -float("inf"), 0.0, -0.0, 4.0]
    b = pa.record_batch({"v": pa.array(vals, pa.float64())})
    cases = [
        (Cmp.Gt, 4.0, [0, 2]),
        (Cmp.Lt, -float("inf"), [1]),
        (Cmp.Eq, float("nan"), [0]),
        (Cmp.Eq, 0.0, [4, 5]),
        (Cmp.Eq, -0.0, [4, 5]),
        (Cmp.Lt, 0.0, [1, 3]),
        (Cmp.GtEq, -0.0, [0, 2, 4, 5, 6]),
    ]
    idx = pa.record_batch({"v": b.column(0), "row": pa.array(np.arange(7), pa.int64())})
    for op, x, rows in cases:
        got, want = run_both(backend, stream, oracle, idx, Predicate.compare(0, op, Lit.float64(x)), [1])
        assert got.column(0).to_pylist() == rows, (op, x)
        assert_batches_equal(got, want)


PREDICATES = [
    Predicate.compare(0, Cmp.GtEq, Lit.int64(0)),
    Predicate.compare(1, Cmp.Lt, Lit.int64(3)),
    Predicate.compare(1, Cmp.Eq, Lit.int32(-4)),
    Predicate.compare(2, Cmp.Gt, Lit.float64(0.5)),
    Predicate.compare(2, Cmp.LtEq, Lit.float32(-0.0)),
    Predicate.compare(3, Cmp.GtEq, Lit.float64(float("nan"))),
    Predicate.compare(3, Cmp.Lt, Lit.float64(-float("inf"))),
    Predicate.compare(3, Cmp.NotEq, Lit.float64(0.0)),
    Predicate.and_(Predicate.compare(0, Cmp.GtEq, Lit.int64(2)), Predicate.compare(3, Cmp.Gt, Lit.float64(4.0))),
    Predicate.or_(Predicate.compare(0, Cmp.Lt, Lit.int64(-40)), Predicate.compare(2, Cmp.Gt, Lit.float64(1.0))),
    plan.or_(plan.and_(Predicate.compare(1, Cmp.Lt, Lit.int64(5)), Predicate.compare(2, Cmp.Gt, Lit.float64(0.5))),
             plan.cmp(plan.mod(col(0), Lit.int64(7)), Cmp.Eq, Lit.int64(3))),
    plan.cmp(plan.add(col(0), col(1)), Cmp.Gt, Lit.int64(10)),
    plan.cmp(plan.mul(col(2), col(3)), Cmp.LtEq, Lit.float64(0.25)),
    plan.cmp(col(0), Cmp.Lt, col(1)),
    plan.cmp(Lit.int64(3), Cmp.Gt, col(0)),
]

PROJECTIONS = [
    [0],
    [4, 0],
    [col(3), plan.add(plan.mul(col(1), Lit.int64(2)), Lit.int64(1)), plan.mul(col(2), col(3))],
    [plan.sub(col(0), col(1)), plan.div(col(3), col(2)), plan.mod(col(1), Lit.int32(3)),
     plan.mul(col(2), Lit.float32(1.5)), plan.add(col(1), Lit.float64(0.25))],
    [plan.div(col(0), Lit.int64(3)), plan.mul(col(0), col(0))],
]


@pytest.mark.parametrize("n", [0, 1, 7, 64, 1000, 4097, 50_000])
@pytest.mark.parametrize("seed", [1, 42])
def test_random_parity(backend, stream, oracle, n, seed, path, small_tiles):
    rng = np.random.default_rng(seed * 1000 + n)
    batch = datagen.random_batch(rng, n)
    for pi, pred in enumerate(PREDICATES):
        for qi, proj in enumerate(PROJECTIONS):
            if (pi + qi) % 3 and n > 1000:
                continue
            got, want = run_both(backend, stream, oracle, batch, pred, proj)
            assert_batches_equal(got, want, names=False)


def test_no_predicate_keeps_every_row(backend, stream, oracle, path):
    rng = np.random.default_rng(3)
    batch = datagen.random_batch(rng, 3000)
    got, want = run_both(backend, stream, oracle, batch, None, [0, 1, 2, 3, 4])
    assert got.num_rows == 3000
    assert_batches_equal(got, want)


def test_sliced_input(backend, stream, oracle, path):
    rng = np.random.default_rng(5)
    batch = datagen.random_batch(rng, 999).slice(13, 700)
    pred = Predicate.compare(2, Cmp.Gt, Lit.float64(0.0))
    got, want = run_both(backend, stream, oracle, batch, pred, [0, 2, 4])
    assert_batches_equal(got, want)


def test_divide_by_zero_is_an_error(backend, stream, path):
    b = pa.record_batch({"x": pa.array([4, 5, 6], pa.int64()), "y": pa.array([1, 0, 2], pa.int64())})
    dev = backend.upload(stream, b)
    with pytest.raises(ExecutionError):
        backend.filter_project(stream, dev, None, [plan.div(col(0), col(1))])
    # a zero divisor in a NULL row is not an error
    b = pa.record_batch({"x": pa.array([4, 5, 6], pa.int64()),
                         "y": pa.array([1, None, 2], pa.int64())})
    b = pa.record_batch({"x": b.column(0), "y": pa.array([1, 0, 2], pa.int64(), mask=np.array([False, True, False]))})
    dev = backend.upload(stream, b)
    out = backend.download(stream, backend.filter_project(stream, dev, None, [plan.div(col(0), col(1))]))
    assert out.column(0).to_pylist() == [4, None, 3]


def test_unsupported_column_is_declined(backend, stream):
    b = pa.record_batch({"s": pa.array(["a", "b"])})
    with pytest.raises(Unsupported) as e:
        backend.upload(stream, b)
    assert e.value.feature == "device.column_type"


def test_vector_null_element_is_declined(backend, stream):
    # columns.rs:211-241: a NULL element inside a valid row cannot travel
    child = pa.array([1.0, None, 3.0, 4.0], pa.float32())
    v = pa.FixedSizeListArray.from_arrays(child, 2)
    with pytest.raises(Unsupported) as e:
        backend.upload(stream, pa.record_batch({"v": v}))
    assert e.value.feature == "device.vector_null_element"
    # under a NULL row it is fine
    v = pa.FixedSizeListArray.from_arrays(child, 2, mask=pa.array([True, False]))
    dev = backend.upload(stream, pa.record_batch({"v": v}))
    assert dev.num_rows == 2


@pytest.mark.parametrize("threshold", [50, 500])
def test_filter_query_parity(backend, stream, oracle, path, threshold):
    # SELECT a, b*2+1, c*d FROM t WHERE b < T AND c > 0.5 OR a % 7 = 3
    n = 300_001
    dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=7), n)
    host = backend.download(stream, dev)
    assert host.column(2).null_count > 0
    pred, proj = datagen.filter_query(threshold)
    got = backend.download(stream, backend.filter_project(stream, dev, pred, proj, ["a", "b2", "cd"]))
    want = oracle.filter_project(host, pred, proj, ["a", "b2", "cd"])
    assert_batches_equal(got, want)
    assert abs(got.num_rows / n - datagen.filter_selectivity(threshold)) < 0.01


@pytest.mark.parametrize("threshold,c_thr,modulus,want", [(30, 0.5, 1000, 0.0098), (180, 0.5, 20, 0.100),
                                                          (470, -1.0, 5, 0.5006)])
def test_filter_query_selectivity_points(backend, stream, oracle, path, threshold, c_thr, modulus, want):
    # the query's shape at about 1%, 10% and 50% selected (the benchmark's
    # three points: T, C and M moved together)
    n = 300_001
    dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=11), n)
    host = backend.download(stream, dev)
    pred, proj = datagen.filter_query(threshold, c_thr, modulus)
    got = backend.download(stream, backend.filter_project(stream, dev, pred, proj, ["a", "b2", "cd"]))
    assert_batches_equal(got, oracle.filter_project(host, pred, proj, ["a", "b2", "cd"]))
    assert abs(got.num_rows / n - want) < 0.01
    assert abs(datagen.filter_selectivity(threshold, c_thr, modulus=modulus) - want) < 0.002


def test_batches_outlive_their_stream(backend, oracle):
    # a batch allocated on a stream may be used and freed after the caller
    # destroyed that stream's handle (the stream lives until its last batch)
    import gc

    s = backend.create_stream()
    dev = backend.upload(s, ref_batch())
    out = backend.filter_project(s, dev, Predicate.compare(0, Cmp.GtEq, Lit.int64(2)), [0, 1])
    del s
    gc.collect()
    s2 = backend.create_stream()
    got = backend.download(s2, out)
    assert got.column(0).to_pylist() == [2, 2, 5]
    del out, dev
    gc.collect()
    s2.synchronize()


@pytest.mark.parametrize("shape", [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2), (4, 1)])
def test_every_launch_shape_is_bit_exact(backend, stream, oracle, shape):
    # the autotuner may pick any instantiated shape: each must give the same rows
    from oxidelake_amd import autotune

    autotune.set_knob("filter.v", shape[0])
    autotune.set_knob("filter.g", shape[1])
    autotune.set_knob("filter.jit", 0)
    try:
        rng = np.random.default_rng(77)
        for n in (1, 1000, 9001):
            batch = datagen.random_batch(rng, n)
            for pred, proj in [(PREDICATES[10], PROJECTIONS[2]), (PREDICATES[5], PROJECTIONS[1]),
                               (PREDICATES[12], PROJECTIONS[3])]:
                got, want = run_both(backend, stream, oracle, batch, pred, proj)
                assert_batches_equal(got, want, names=False)
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=5), 100_003)
        host = backend.download(stream, dev)
        pred, proj = datagen.filter_query(50)
        got = backend.download(stream, backend.filter_project(stream, dev, pred, proj))
        assert_batches_equal(got, oracle.filter_project(host, pred, proj), names=False)
    finally:
        autotune.set_knob("filter.v", -1)
        autotune.set_knob("filter.g", -1)
        autotune.set_knob("filter.jit", -1)


@pytest.mark.parametrize("persistent", [1, 0])
@pytest.mark.parametrize("rows", [1, 2, 4, 6, 8, 16, 20, 24, 32])
def test_every_rows_per_lane_is_bit_exact(backend, stream, oracle, rows, persistent):
    # generated kernels with S rows per lane (tiles of 256 * S rows); above
    # 16 the tile scan covers two (step, wave) counts per lane
    autotune.set_knob("filter.rows", rows)
    autotune.set_knob("filter.persistent", persistent)
    try:
        rng = np.random.default_rng(78)
        for n in (1, 1000, 9001):
            batch = datagen.random_batch(rng, n)
            for pred, proj in [(PREDICATES[10], PROJECTIONS[2]), (PREDICATES[12], PROJECTIONS[3])]:
                got, want = run_both(backend, stream, oracle, batch, pred, proj)
                assert_batches_equal(got, want, names=False)
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=5), 100_003)
        host = backend.download(stream, dev)
        pred, proj = datagen.filter_query(50)
        outs = backend.filter_project_batches(stream, dev, 8192, pred, proj)
        for i, o in enumerate(outs):
            want = oracle.filter_project(host.slice(i * 8192, 8192), pred, proj)
            assert_batches_equal(backend.download(stream, o), want, names=False)
    finally:
        autotune.set_knob("filter.rows", -1)
        autotune.set_knob("filter.persistent", -1)


@pytest.mark.parametrize("loads,stores", [(0, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("rows", [16, 24])
def test_nontemporal_loads_and_stores_are_bit_exact(backend, stream, oracle, rows, loads, stores):
    # knobs jit.stream_loads / jit.stream_stores (the autotuner may pick
    # nontemporal projection stores): the same bits either way
    autotune.set_knob("filter.rows", rows)
    autotune.set_knob("jit.stream_loads", loads)
    autotune.set_knob("jit.stream_stores", stores)
    try:
        rng = np.random.default_rng(90 + rows)
        for n in (1, 1000, 20_001):
            batch = datagen.random_batch(rng, n)
            for pred, proj in [(PREDICATES[10], PROJECTIONS[2]), (PREDICATES[12], PROJECTIONS[3])]:
                got, want = run_both(backend, stream, oracle, batch, pred, proj)
                assert_batches_equal(got, want, names=False)
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=7), 300_001)
        host = backend.download(stream, dev)
        for t in (5, 50, 470):
            pred, proj = datagen.filter_query(t)
            outs = backend.filter_project_batches(stream, dev, 65_536, pred, proj)
            for i, o in enumerate(outs):
                want = oracle.filter_project(host.slice(i * 65_536, 65_536), pred, proj)
                assert_batches_equal(backend.download(stream, o), want, names=False)
    finally:
        for k in ("filter.rows", "jit.stream_loads", "jit.stream_stores"):
            autotune.set_knob(k, -1)


STAGED_KNOBS = ("filter.staged", "filter.chunk_rows", "filter.chunks", "filter.stage_rows", "filter.staged_waves",
                "filter.stage_nt")


@pytest.mark.parametrize("chunk_rows,chunks,stage_rows,waves,nt", [
    (4, 8, 3072, 8, 1),   # the default staged shape
    (4, 8, 3072, 8, 0),   # its copy out with plain stores
    (8, 3, 2048, 4, 1),
    (4, 6, 2048, 4, 1),
    (12, 2, 1024, 4, 1),
    (16, 2, 4096, 4, 1),
    (8, 3, 64, 4, 1),     # a 64-row stage: most tiles store their later chunks directly
    (8, 2, 8, 8, 1),      # 8 waves, an 8-row stage
    (1, 1, 8, 4, 1),
    (3, 5, 16, 16, 0),
])
def test_staged_kernel_is_bit_exact(backend, stream, oracle, chunk_rows, chunks, stage_rows, waves, nt):
    # kept rows' projections parked in LDS, rows loaded in chunks; tiles that
    # keep more rows than the stage holds store their later chunks directly
    # (filter.staged = 2: the staged kernel whatever the kept fraction)
    for k, v in zip(STAGED_KNOBS, (2, chunk_rows, chunks, stage_rows, waves, nt)):
        autotune.set_knob(k, v)
    try:
        rng = np.random.default_rng(70 + chunk_rows * chunks)
        for n in (1, 1000, 30_001):
            batch = datagen.random_batch(rng, n)
            for pred, proj in [(PREDICATES[10], PROJECTIONS[2]), (PREDICATES[12], PROJECTIONS[1]),
                               (None, PROJECTIONS[3]), (PREDICATES[0], [4, 0])]:
                got, want = run_both(backend, stream, oracle, batch, pred, proj)
                assert_batches_equal(got, want, names=False)
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=5), 1_000_003)
        host = backend.download(stream, dev)
        for t in (5, 50, 470):
            pred, proj = datagen.filter_query(t)
            for batch_rows in (65_536, 1_000_000):
                outs = backend.filter_project_batches(stream, dev, batch_rows, pred, proj)
                for i, o in enumerate(outs):
                    want = oracle.filter_project(host.slice(i * batch_rows, batch_rows), pred, proj)
                    assert_batches_equal(backend.download(stream, o), want, names=False)
        # a zero divisor in the last row of a large batch: with every row
        # kept, that row's chunk is past any stage and stored directly
        n = 200_000
        y = np.ones(n, np.int64)
        y[-1] = 0
        b = pa.record_batch({"x": pa.array(np.arange(n, dtype=np.int64)), "y": pa.array(y)})
        dev = backend.upload(stream, b)
        with pytest.raises(ExecutionError):
            backend.filter_project(stream, dev, None, [plan.div(col(0), col(1))])
    finally:
        for k in STAGED_KNOBS:
            autotune.set_knob(k, -1)


@pytest.mark.parametrize("waves,rows", [(8, 16), (12, 16), (16, 8), (8, 24), (12, 3), (16, 1)])
def test_every_waves_per_workgroup_is_bit_exact(backend, stream, oracle, waves, rows):
    # tiles of 64 * W * S rows; S * W above 64 (step, wave) counts scans two
    # or more counts per lane
    autotune.set_knob("filter.rows", rows)
    autotune.set_knob("filter.waves", waves)
    try:
        rng = np.random.default_rng(80 + waves)
        for n in (1, 1000, 30_001):
            batch = datagen.random_batch(rng, n)
            for pred, proj in [(PREDICATES[10], PROJECTIONS[2]), (PREDICATES[12], PROJECTIONS[1])]:
                got, want = run_both(backend, stream, oracle, batch, pred, proj)
                assert_batches_equal(got, want, names=False)
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=6), 1_000_003)
        host = backend.download(stream, dev)
        pred, proj = datagen.filter_query(50)
        outs = backend.filter_project_batches(stream, dev, 65_536, pred, proj)
        for i, o in enumerate(outs):
            want = oracle.filter_project(host.slice(i * 65_536, 65_536), pred, proj)
            assert_batches_equal(backend.download(stream, o), want, names=False)
    finally:
        autotune.set_knob("filter.rows", -1)
        autotune.set_knob("filter.waves", -1)


@pytest.mark.parametrize("rows,batch_rows", [(4, 65_536), (8, 1_000_000), (2, 4_096)])
def test_persistent_grid_takes_many_tiles_per_workgroup(backend, stream, oracle, rows, batch_rows):
    # one workgroup per CU: every workgroup compacts several tiles in turn,
    # across batch boundaries, with the next tile's loads in flight
    autotune.set_knob("filter.rows", rows)
    autotune.set_knob("filter.persistent", 1)
    autotune.set_knob("filter.grid", 1)
    try:
        n = 2_000_003
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=11), n)
        host = backend.download(stream, dev)
        pred, proj = datagen.filter_query(300)
        outs = backend.filter_project_batches(stream, dev, batch_rows, pred, proj)
        assert len(outs) == -(-n // batch_rows)
        for i, o in enumerate(outs):
            want = oracle.filter_project(host.slice(i * batch_rows, batch_rows), pred, proj)
            assert_batches_equal(backend.download(stream, o), want, names=False)
    finally:
        for k in ("filter.rows", "filter.persistent", "filter.grid"):
            autotune.set_knob(k, -1)


@pytest.mark.parametrize("rows,waves,batch_rows,prepared", [(8, 8, 300_000, 0), (4, 4, 4_096, 0),
                                                          (16, 4, 65_536, 1), (8, 8, 32, 1)])
def test_batch_per_workgroup_is_bit_exact(backend, stream, oracle, rows, waves, batch_rows, prepared):
    # one workgroup per batch (filter.wg = 2: whatever the batch count),
    # walking its batches' tiles in order with a running offset; more
    # batches than workgroups at the smaller batch sizes, a ragged last
    # batch, and the query at its 1% and 50% selectivity points
    autotune.set_knob("filter.wg", 2)
    autotune.set_knob("filter.wg_rows", rows)
    autotune.set_knob("filter.wg_waves", waves)
    autotune.set_knob("jit.params", prepared)
    try:
        n = 700_003
        dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=13), n)
        host = backend.download(stream, dev)
        for b_thr, c_thr, modulus in [(50, 0.5, 7), (30, 0.5, 1000), (470, -1.0, 5)]:
            pred, proj = datagen.filter_query(b_thr, c_thr, modulus)
            outs = backend.filter_project_batches(stream, dev, batch_rows, pred, proj)
            assert len(outs) == -(-n // batch_rows)
            check = range(len(outs)) if len(outs) <= 64 else list(range(0, len(outs), 97)) + [len(outs) - 1]
            for i in check:
                want = oracle.filter_project(host.slice(i * batch_rows, batch_rows), pred, proj)
                assert_batches_equal(backend.download(stream, outs[i]), want, names=False)
    finally:
        for k in ("filter.wg", "filter.wg_rows", "filter.wg_waves", "jit.params"):
            autotune.set_knob(k, -1)


def test_generated_kernel_is_the_one_that_runs(backend, stream):
    # the per-query kernel, not the interpreter, runs a query in its grammar
    from oxidelake_amd._capi import lib

    dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=3), 10_000)
    pred, proj = datagen.filter_query(50)
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        backend.filter_project(stream, dev, pred, proj)
        assert autotune.launches("filter_project") == 1
        assert autotune.launches("filter_project_vm") == 0
        autotune.set_knob("filter.jit", 0)
        backend.filter_project(stream, dev, pred, proj)
        assert autotune.launches("filter_project_vm") == 1
    finally:
        autotune.set_knob("filter.jit", -1)
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()


def test_zero_divisor_in_a_dropped_row_is_no_error(backend, stream, oracle, path):
    # projections are evaluated over the kept rows only
    x = np.arange(64, dtype=np.int64)
    y = np.where(x % 4 == 1, 0, 3).astype(np.int64)
    b = pa.record_batch({"x": pa.array(x), "y": pa.array(y)})
    pred = Predicate.compare(1, Cmp.NotEq, Lit.int64(0))
    for shape in [(1, 1), (4, 1), (2, 2)]:
        autotune.set_knob("filter.v", shape[0])
        autotune.set_knob("filter.g", shape[1])
        try:
            got, want = run_both(backend, stream, oracle, b, pred, [plan.div(col(0), col(1))])
        finally:
            autotune.set_knob("filter.v", -1)
            autotune.set_knob("filter.g", -1)
        assert_batches_equal(got, want, names=False)


@pytest.mark.parametrize("batch_rows", [32, 4096, 65_536])
def test_many_batches_in_one_launch(backend, stream, oracle, path, batch_rows):
    # each batch of the input compacted on its own, as the reference filters
    # batch by batch (exec/filter.rs:182-210)
    n = 250_003
    dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=9), n)
    host = backend.download(stream, dev)
    pred, proj = datagen.filter_query(500)
    outs = backend.filter_project_batches(stream, dev, batch_rows, pred, proj, ["a", "b2", "cd"])
    assert len(outs) == -(-n // batch_rows)
    for i, o in enumerate(outs):
        got = backend.download(stream, o)
        want = oracle.filter_project(host.slice(i * batch_rows, batch_rows), pred, proj, ["a", "b2", "cd"])
        assert_batches_equal(got, want)
    # views share buffers: freeing some leaves the others intact
    keep = outs[-1]
    del outs
    import gc

    gc.collect()
    got = backend.download(stream, keep)
    last = (n - 1) // batch_rows
    assert_batches_equal(got, oracle.filter_project(host.slice(last * batch_rows, batch_rows), pred, proj,
                                                    ["a", "b2", "cd"]))


def test_prepared_filter_and_batch_sets(backend, stream, path):
    # a plan encoded once runs like the per-call form; a multi-batch result
    # frees its batches together, after its last view is gone
    from oxidelake_amd import datagen

    dev = datagen.table(backend, stream, datagen.filter_table_columns(seed=5), 300_000)
    pred, proj = datagen.filter_query(120)
    prep = backend.prepare_filter(pred, proj, ["a", "b2", "cd"])
    want = backend.filter_project_batches(stream, dev, 65_536, pred, proj, ["a", "b2", "cd"])
    got = prep.run_batches(stream, dev, 65_536)
    assert len(got) == len(want) == 5
    for g, w in zip(got, want):
        assert backend.download(stream, g).equals(backend.download(stream, w))
    one = prep.run(stream, dev)
    assert one.num_rows == sum(b.num_rows for b in got)
    last = got[-1]
    del got, want
    assert last.num_rows >= 0  # the view keeps its set alive
    assert [b.num_rows for b in prep.run_batches(stream, dev, 65_536)[1:3]] == [
        b.num_rows for b in backend.filter_project_batches(stream, dev, 65_536, pred, proj)[1:3]]


@pytest.mark.parametrize("lanes", [1, 3])
def test_streaming_host_batches(backend, stream, oracle, lanes):
    # the plain drop-in path over host batches (upload, filter, download per
    # batch, on `lanes` streams): every batch as the oracle filters it
    from oxidelake_amd import datagen
    from oxidelake_amd.streaming import StreamingFilter

    rng = np.random.default_rng(31)
    batches = [datagen.random_batch(rng, n) for n in (1, 5000, 0, 70_001, 333)]
    pred = plan.or_(plan.cmp(col(1), Cmp.Lt, Lit.int32(0)), plan.cmp(col(2), Cmp.Gt, Lit.float64(0.25)))
    proj = [0, plan.add(col(1), Lit.int64(3)), 3]
    got = StreamingFilter(backend, pred, proj, lanes=lanes).run(batches)
    for b, g in zip(batches, got):
        assert_batches_equal(g, oracle.filter_project(b, pred, proj), names=False)


def test_operator_sub_batch_and_host_columns(backend, stream, oracle):
    # only the columns the predicate and computed projections read travel;
    # a projected string column stays on the host and is gathered by row id;
    # a predicate over a column the device does not take falls back
    from oxidelake_amd.streaming import FilterOperator

    rng = np.random.default_rng(12)
    base = datagen.random_batch(rng, 20_000).select(["i", "j", "f", "g"])
    tags = pa.array([f"r{k}" for k in range(base.num_rows)])
    small = pa.array(rng.integers(0, 5, base.num_rows).astype(np.int8))
    batch = pa.RecordBatch.from_arrays(base.columns + [tags, small], names=base.schema.names + ["tag", "small"])
    pred = plan.cmp(col(0), Cmp.Gt, Lit.int64(10))
    proj = [4, plan.mul(col(1), Lit.int64(2)), 3]
    op = FilterOperator(backend, pred, proj, names=["tag", "j2", "g"])
    got = op.process(batch)
    numeric = batch.select(["i", "j", "f", "g"])
    want = oracle.filter_project(numeric, pred, [plan.mul(col(1), Lit.int64(2)), 3])
    kept = np.nonzero(np.nan_to_num(numeric.column(0).to_numpy(zero_copy_only=False), nan=-1) > 10)[0]
    assert got.num_rows == want.num_rows == len(kept)
    assert got.column(0).to_pylist() == tags.take(pa.array(kept)).to_pylist()
    assert_batches_equal(got.select(["j2", "g"]), want, names=False)
    assert (op.device_batches, op.fallback_batches) == (1, 0)
    # a predicate on the Int8 column: the CPU path runs the batch
    marker = batch.slice(0, 0)
    op2 = FilterOperator(backend, plan.cmp(col(5), Cmp.Eq, Lit.int64(1)), [0], cpu_fallback=lambda b: marker)
    assert op2.process(batch) is marker and (op2.device_batches, op2.fallback_batches) == (0, 1)

"""Generated filter kernels, on the CPU: hipRTC compiles the kernel of every
query shape the GPU parity tests run (no device is needed to compile), and
the generated source has the structure the interpreter's compiler implies.
"""
import pytest

from oxidelake_amd import autotune, backend, datagen, plan
from oxidelake_amd.backend import PlanError, Unsupported
from oxidelake_amd.plan import Comparison as Cmp
from oxidelake_amd.plan import Lit, Predicate, col
from tests.test_filter_gpu import PREDICATES, PROJECTIONS

# random_batch's schema: i Int64, j Int32, f Float32, g Float64, e vector<3>
RANDOM = [(plan.INT64, True), (plan.INT32, True), (plan.FLOAT32, True), (plan.FLOAT64, True),
          (plan.F32_VECTOR, 3, True)]
FILTER_TABLE = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT32, True), (plan.FLOAT64, False)]


@pytest.fixture(autouse=True)
def four_rows():
    autotune.set_knob("filter.rows", 4)
    yield
    autotune.set_knob("filter.rows", -1)


@pytest.mark.parametrize("pi", range(len(PREDICATES)))
def test_every_test_predicate_compiles(pi):
    src = backend.filter_kernel_source(RANDOM, PREDICATES[pi], [0], compile=True)
    assert "oxr_filter_jit" in src


@pytest.mark.parametrize("qi", range(len(PROJECTIONS)))
def test_every_test_projection_compiles(qi):
    backend.filter_kernel_source(RANDOM, None, PROJECTIONS[qi], compile=True)


def test_filter_query_kernel_shape():
    autotune.set_knob("filter.rows", -1)
    pred, proj = datagen.filter_query()
    src = backend.filter_kernel_source(FILTER_TABLE, pred, proj, compile=True)
    # four 1-2 register columns plus one validity word: 16 rows per lane (8
    # on the persistent kernel, which holds two tiles' rows)
    assert "filter_tile<Q, 16, 4>" in src
    autotune.set_knob("filter.persistent", 1)
    try:
        assert "filter_persistent<Q, 8, 4>" in backend.filter_kernel_source(FILTER_TABLE, pred, proj, compile=True)
    finally:
        autotune.set_knob("filter.persistent", -1)
    # a % 7 by a constant divisor (the compiler's multiply-high), no error path
    assert "(x.c0 % ((int64_t)0x0000000000000007ull))" in src
    # literals are constants: b < 50, then c > 0.5 by its bit pattern
    assert "t3 < ((int64_t)0x0000000000000032ull)" in src
    assert "order_key64(as_f64(0x3fe0000000000000ull))" in src
    pred2, _ = datagen.filter_query(500)
    src2 = backend.filter_kernel_source(FILTER_TABLE, pred2, proj)
    assert src2 == src.replace("0x0000000000000032ull", "0x00000000000001f4ull")
    # only c*d can be NULL
    assert "kValidOuts = 1;" in src


def test_division_by_zero_or_minus_one_keeps_its_error_path():
    c = [(plan.INT64, False)]
    for d, fn in [(0, "i64_divmod(false"), (-1, "i64_divmod(false"), (3, " / ((int64_t)")]:
        src = backend.filter_kernel_source(c, None, [plan.div(col(0), Lit.int64(d))], compile=d == -1)
        assert fn in src, d
    src = backend.filter_kernel_source([(plan.INT32, True)], None, [plan.mod(col(0), Lit.int32(-1))])
    assert "i32_divmod(true" in src


def test_nullability_changes_the_kernel():
    pred, proj = datagen.filter_query()
    all_valid = [(t, False) for t, _ in FILTER_TABLE]
    a = backend.filter_kernel_source(all_valid, pred, proj)
    b = backend.filter_kernel_source(FILTER_TABLE, pred, proj)
    assert a != b and "row_bit" not in a


def test_vector_projection_is_a_wave_copy():
    src = backend.filter_kernel_source(RANDOM, PREDICATES[0], [4, 0], compile=True)
    assert "copy_vector_rows" in src and "kVecOuts = 1;" in src


@pytest.mark.parametrize("persistent", [1, 0])
@pytest.mark.parametrize("rows", [1, 2, 8, 16, 24])
def test_every_rows_per_lane_compiles(rows, persistent):
    autotune.set_knob("filter.rows", rows)
    autotune.set_knob("filter.persistent", persistent)
    try:
        pred, proj = datagen.filter_query()
        src = backend.filter_kernel_source(FILTER_TABLE, pred, proj, compile=True)
        assert f"{'filter_persistent' if persistent else 'filter_tile'}<Q, {rows}, 4>" in src
    finally:
        autotune.set_knob("filter.persistent", -1)


@pytest.mark.parametrize("chunk_rows,chunks,stage_rows", [(8, 3, 2048), (4, 6, 1536), (16, 2, 4096), (1, 1, 8)])
def test_staged_kernel_compiles(chunk_rows, chunks, stage_rows):
    knobs = {"filter.staged": 1, "filter.chunk_rows": chunk_rows, "filter.chunks": chunks,
             "filter.stage_rows": stage_rows, "filter.staged_waves": 4}
    for k, v in knobs.items():
        autotune.set_knob(k, v)
    try:
        pred, proj = datagen.filter_query()
        src = backend.filter_kernel_source(FILTER_TABLE, pred, proj, compile=True)
        assert f"filter_staged<Q, {chunk_rows}, {chunks}, 4, {stage_rows}>" in src
        # beside the tile kernel, which keeps its own shape
        assert "filter_tile<Q, 4, 4>" in src
        # three 8-byte projections per kept row in the stage, copied out
        # with nontemporal stores (knob filter.stage_nt, default 1)
        assert "kStageBytes = 24;" in src
        assert "st_stream(&o.o2[i], t[i]);" in src
        # vector projections keep the tile kernel's whole-wave copies
        src = backend.filter_kernel_source(RANDOM, PREDICATES[0], [4, 0], compile=True)
        assert "filter_tile<Q, 4, 4>" in src and "filter_staged<" not in src
    finally:
        for k in knobs:
            autotune.set_knob(k, -1)


@pytest.mark.parametrize("knobs", [
    {"filter.chunk_rows": 8, "filter.staged_waves": 16},  # 128 (step, wave) counts in a chunk
    {"filter.chunk_rows": 16, "filter.chunks": 3},   # 48 rows per lane
    {"filter.stage_rows": 8192},                     # 192 KiB of stage
])
def test_staged_shapes_outside_the_kernel_are_refused(knobs):
    autotune.set_knob("filter.staged", 1)
    for k, v in knobs.items():
        autotune.set_knob(k, v)
    try:
        pred, proj = datagen.filter_query()
        with pytest.raises(PlanError):
            backend.filter_kernel_source(FILTER_TABLE, pred, proj)
    finally:
        for k in ("filter.staged", *knobs):
            autotune.set_knob(k, -1)


def test_spill_records_read_nontemporal_compiles():
    # knob agg.spill_nt (default 1): the scatter and the fold read the
    # spill's records with nontemporal loads (4-word records, uint4 loads)
    c3 = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT64, False)]
    assert "kSpillNT = true;" in backend.aggregate_kernel_source(c3, datagen.star_query_spec(), probe_key=0)
    autotune.set_knob("agg.spill_nt", 0)
    try:
        src = backend.aggregate_kernel_source(c3, datagen.star_query_spec(), probe_key=0, compile=True)
        assert "kSpillNT = false;" in src
    finally:
        autotune.set_knob("agg.spill_nt", -1)



    with pytest.raises(PlanError):
        backend.filter_kernel_source(FILTER_TABLE, Predicate.compare(9, Cmp.Eq, Lit.int64(1)), [0])
    with pytest.raises(Unsupported):
        backend.filter_kernel_source(RANDOM, Predicate.compare(4, Cmp.Eq, Lit.int64(1)), [0])
    with pytest.raises(Unsupported):
        backend.filter_kernel_source(FILTER_TABLE, None, [plan.cmp(col(0), Cmp.Eq, col(1))])


# ------------------------------------------------------------- aggregation
from tests.test_aggregate_gpu import SPECS as AGG_SPECS  # noqa: E402

AGG_COLS = [(plan.INT64, True), (plan.INT32, True), (plan.FLOAT32, True), (plan.FLOAT64, True)]


@pytest.mark.parametrize("si", range(len(AGG_SPECS)))
def test_every_test_aggregation_compiles(si):
    src = backend.aggregate_kernel_source(AGG_COLS, AGG_SPECS[si], compile=True)
    assert "oxr_aggregate_jit" in src


def test_groupby_and_star_kernels():
    autotune.set_knob("agg.rows", -1)
    c1 = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT64, False), (plan.F32_VECTOR, 8, False)]
    src = backend.aggregate_kernel_source(c1, datagen.groupby_query_spec(), compile=True)
    # k and v only (the id and vector columns are never loaded), no NULLs
    assert "x.c1" in src and "x.c2" in src and "x.c0" not in src and "x.c3" not in src
    assert "kAnyNullable = false" in src and "aggregate_tiles<Q, 16>" in src
    c3 = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT64, False)]
    src = backend.aggregate_kernel_source(c3, datagen.star_query_spec(), probe_key=0, compile=True)
    assert "join_group(p, (int64_t)x.c0, key, kv, hid)" in src
    # spill records of 4 words: group, SUM(qty) (Int32), SUM(qty*price) (FP64), COUNT(*) (none)
    assert "kSpillWords[kAggs] = {1, 2, 0}" in src
    assert "spill_scatter<Q>(a)" in src and "spill_fold<Q>(a)" in src


@pytest.mark.parametrize("exact", [0, 1])
def test_keyless_kernel_only_for_global_aggregates(exact):
    autotune.set_knob("agg.exact_f64", exact)
    try:
        glob = [s for s in AGG_SPECS if s.group_by < 0]
        keyed = [s for s in AGG_SPECS if s.group_by >= 0]
        assert glob and keyed
        for spec in glob:
            src = backend.aggregate_kernel_source(AGG_COLS, spec, compile=True)
            assert "oxr_aggregate_keyless" in src and "aggregate_keyless<Q, " in src
        for spec in keyed:
            assert "oxr_aggregate_keyless" not in backend.aggregate_kernel_source(AGG_COLS, spec)
        c3 = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT64, False)]
        assert "keyless" not in backend.aggregate_kernel_source(c3, datagen.star_query_spec(), probe_key=0)
    finally:
        autotune.set_knob("agg.exact_f64", -1)


def test_aggregation_without_aggregates_uses_the_interpreter():
    from oxidelake_amd.backend import Unsupported

    with pytest.raises(Unsupported):
        backend.aggregate_kernel_source(AGG_COLS, plan.AggregateSpec(0, []))


def test_prepared_mode_shares_one_kernel_across_literals():
    # knob jit.params: literals become kernel arguments (one binary per shape)
    autotune.set_knob("jit.params", 1)
    try:
        p50, proj = datagen.filter_query(50)
        p500, _ = datagen.filter_query(500)
        a = backend.filter_kernel_source(FILTER_TABLE, p50, proj, compile=True)
        assert a == backend.filter_kernel_source(FILTER_TABLE, p500, proj) and "L[0]" in a
        # no literal-divisor shortcut: the checked division path
        assert "i64_divmod(true" in a
        spec = lambda t: plan.AggregateSpec(1, [plan.Aggregate(plan.AggregateFunction.Sum, col(0))],
                                            filter=Predicate.compare(1, Cmp.Lt, Lit.int64(t)))
        x = backend.aggregate_kernel_source(FILTER_TABLE, spec(10), compile=True)
        assert x == backend.aggregate_kernel_source(FILTER_TABLE, spec(400)) and "L[0]" in x
    finally:
        autotune.set_knob("jit.params", -1)

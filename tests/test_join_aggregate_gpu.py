"""GPU parity: fused inner join + grouped aggregation (oxr_agg_update_join)
vs the oracle's join (crates/oxidelake-device/src/cpu/kernels.rs:270-472)
followed by its aggregation (kernels.rs:478-818) over the joined rows,
grouped by a build-side column. Integers and keys bit-exact; Float64 sums
within 1e-12 relative of the oracle's compensated sum.
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import datagen, plan
from oxidelake_amd.backend import AggState, PlanError
from oxidelake_amd.plan import AggregateFunction as F
from oxidelake_amd.plan import Aggregate, AggregateSpec, Lit, col
from oxidelake_amd.plan import Comparison as Cmp
from tests.test_aggregate_gpu import assert_agg_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["generated", "exact", "spill", "hot", "interpreter"])
def path(request):
    # unique-key joins run on the generated kernel, others on the
    # interpreter; exact: the generated kernel with compensated FP64 sums;
    # spill: cold groups spilled, partitioned and folded in LDS; hot: the
    # spill plus sampled hot groups in dense LDS accumulators
    from oxidelake_amd import autotune
    from tests.test_aggregate_gpu import HOT_KNOBS, SPILL_KNOBS

    autotune.set_knob("agg.jit", 0 if request.param == "interpreter" else 1)
    autotune.set_knob("agg.exact_f64", 1 if request.param == "exact" else 0)
    if request.param in ("spill", "hot"):
        for k, v in (SPILL_KNOBS if request.param == "spill" else HOT_KNOBS).items():
            autotune.set_knob(k, v)
    yield request.param
    for k in ("agg.jit", "agg.exact_f64", *HOT_KNOBS):
        autotune.set_knob(k, -1)


# fact: fk, qty, price; dim: pk, grp
AGGS = [Aggregate(F.Sum, col(1)), Aggregate(F.Sum, plan.mul(col(1), col(2))), Aggregate(F.CountStar, None),
        Aggregate(F.Min, col(2)), Aggregate(F.Max, col(1)), Aggregate(F.Count, col(1))]


def tables(rng, nf, nd, groups, unique=True, nulls=True):
    pk = rng.permutation(nd).astype(np.int64) if unique else rng.integers(0, nd // 2 + 1, nd).astype(np.int64)
    grp = rng.integers(0, groups, nd).astype(np.int32)
    fk = rng.integers(-5, nd + 5, nf).astype(np.int64)  # some keys miss the dimension
    qty = rng.integers(1, 101, nf).astype(np.int32)
    price = rng.uniform(1, 1000, nf)
    m = (lambda p, n: rng.random(n) < p) if nulls else (lambda p, n: np.zeros(n, bool))
    fact = pa.record_batch({"fk": pa.array(fk, mask=m(0.05, nf)), "qty": pa.array(qty, mask=m(0.05, nf)),
                            "price": pa.array(price, mask=m(0.05, nf))})
    dim = pa.record_batch({"pk": pa.array(pk, mask=m(0.02, nd)), "grp": pa.array(grp, mask=m(0.03, nd))})
    return fact, dim


def oracle_join_agg(oracle, fact, dim, aggs, filt=None):
    joined = oracle.hash_join(fact, dim, 0, 0)  # fact columns, then dim's (pk 3, grp 4)
    spec = AggregateSpec(4, aggs, filter=filt)
    return oracle.aggregate(joined, spec, sum_mode=1), spec


def device_join_agg(backend, stream, fact, dim, aggs, filt=None, pieces=1):
    dfact = backend.upload(stream, fact)
    ddim = backend.upload(stream, dim)
    table = backend.prepare_join_build(stream, ddim, 0)
    st = AggState(backend, stream, dfact, AggregateSpec(-1, aggs, filter=filt))
    n = fact.num_rows
    step = -(-n // pieces) if n else 1
    for i in range(pieces):
        part = dfact if pieces == 1 else backend.upload(stream, fact.slice(i * step, step))
        st.update_join(stream, table, ddim, 1, part, 0)
    return backend.download(stream, st.finish(stream))


@pytest.mark.parametrize("nf,nd,groups", [(1, 1, 1), (1000, 50, 7), (50_000, 3000, 200), (200_001, 20_000, 5000)])
@pytest.mark.parametrize("unique", [True, False])
def test_join_aggregate_parity(backend, stream, oracle, nf, nd, groups, unique):
    rng = np.random.default_rng(nf + nd + unique)
    fact, dim = tables(rng, nf, nd, groups, unique)
    want, spec = oracle_join_agg(oracle, fact, dim, AGGS)
    got = device_join_agg(backend, stream, fact, dim, AGGS)
    assert got.schema.names[0] == "grp"
    assert_agg_equal(got, want, spec)


def test_filter_and_streaming_pieces(backend, stream, oracle):
    rng = np.random.default_rng(3)
    fact, dim = tables(rng, 60_000, 4000, 300)
    filt = plan.cmp(col(1), Cmp.Gt, Lit.int32(40))
    want, spec = oracle_join_agg(oracle, fact, dim, AGGS, filt)
    got = device_join_agg(backend, stream, fact, dim, AGGS, filt, pieces=4)
    assert_agg_equal(got, want, spec)


def test_star_query_with_zipf_skew(backend, stream, oracle, path):
    # SUM(qty), SUM(qty*price), COUNT(*) GROUP BY grp over fact JOIN dim ON fk = pk
    nf, nd, groups = 2_000_000, 100_000, 10_000
    fact_cols, dim_cols = datagen.star_columns(nf, nd, groups, seed=5)
    dfact = datagen.table(backend, stream, fact_cols, nf)
    ddim = datagen.table(backend, stream, dim_cols, nd)
    fact, dim = backend.download(stream, dfact), backend.download(stream, ddim)
    aggs = datagen.star_query_spec().aggregates
    table = backend.prepare_join_build(stream, ddim, 0)
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    st = AggState(backend, stream, dfact, AggregateSpec(-1, aggs))
    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        st.update_join(stream, table, ddim, 1, dfact, 0)
        hot = autotune.launches("aggregate_hot")
    finally:
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    got = backend.download(stream, st.finish(stream))
    want, spec = oracle_join_agg(oracle, fact, dim, aggs)
    assert got.num_rows == want.num_rows <= groups
    assert_agg_equal(got, want, spec)
    assert hot == (1 if path == "hot" else 0)
    # Zipf: the hottest key alone is a large share of the rows
    counts = got.column(3).to_numpy()
    assert counts.sum() == nf and counts.max() > nf // 20


def test_hot_map_is_kept_with_the_probe_batch(backend, stream, oracle, path):
    """The hot-group map of a probe batch through a join table is built by
    the first update and taken by every later one, whatever its filter or
    stream; another join table builds its own; knob agg.hot_cache 0 builds
    one per update. Every result against the oracle."""
    if path != "hot":
        pytest.skip("hot-group kernel only")
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    # (groups past 4x the 1,536 hot slots: the hot kernel's condition)
    nf, nd, groups = 1_000_000, 100_000, 10_000
    fact_cols, dim_cols = datagen.star_columns(nf, nd, groups, seed=9)
    dfact = datagen.table(backend, stream, fact_cols, nf)
    ddim = datagen.table(backend, stream, dim_cols, nd)
    fact, dim = backend.download(stream, dfact), backend.download(stream, ddim)
    aggs = datagen.star_query_spec().aggregates
    filt = plan.cmp(col(1), Cmp.Gt, Lit.int32(50))
    table = backend.prepare_join_build(stream, ddim, 0)
    other_stream = backend.create_stream()

    def run(tbl, f, strm):
        st = AggState(backend, strm, dfact, AggregateSpec(-1, aggs, filter=f))
        st.update_join(strm, tbl, ddim, 1, dfact, 0)
        return backend.download(strm, st.finish(strm))

    lib.oxr_profile_enable(1)
    lib.oxr_profile_reset()
    try:
        got = [(run(table, None, stream), None), (run(table, filt, stream), filt),
               (run(table, None, other_stream), None)]
        assert autotune.launches("aggregate_hot_map") == 1
        assert autotune.launches("aggregate_hot") == 3
        table2 = backend.prepare_join_build(stream, ddim, 0)
        got.append((run(table2, filt, stream), filt))
        assert autotune.launches("aggregate_hot_map") == 2
        autotune.set_knob("agg.hot_cache", 0)
        got.append((run(table, None, stream), None))
        got.append((run(table, None, stream), None))
        assert autotune.launches("aggregate_hot_map") == 4
    finally:
        autotune.set_knob("agg.hot_cache", -1)
        lib.oxr_profile_enable(0)
        lib.oxr_profile_reset()
    for g, f in got:
        want, spec = oracle_join_agg(oracle, fact, dim, aggs, f)
        assert_agg_equal(g, want, spec)


def test_plain_and_joined_updates_do_not_mix(backend, stream):
    rng = np.random.default_rng(4)
    fact, dim = tables(rng, 100, 10, 3, nulls=False)
    dfact, ddim = backend.upload(stream, fact), backend.upload(stream, dim)
    table = backend.prepare_join_build(stream, ddim, 0)
    st = AggState(backend, stream, dfact, AggregateSpec(-1, AGGS))
    st.update_join(stream, table, ddim, 1, dfact, 0)
    with pytest.raises(PlanError):
        st.update(stream, dfact)
    # the build batch must be the one the table indexes
    other = backend.upload(stream, dim.slice(0, 5))
    with pytest.raises(PlanError):
        st.update_join(stream, table, other, 1, dfact, 0)

"""The serving loop on the device: every request's result is the same when
it is served in the mix (batched with other vector queries or filter_agg
requests, beside other request kinds on other streams) as when it runs
alone through the plain path; and the aggregate
requests agree with the oracle (integers exact, Float64 sums within 1e-12
relative of its compensated sum)."""
import copy

import numpy as np
import pytest

from oxidelake_amd import datagen, plan, serving
from oxidelake_amd.backend import VEC_BF16, VectorIndex
from oxidelake_amd.plan import AggregateSpec
from tests.test_aggregate_gpu import assert_agg_equal

pytestmark = pytest.mark.gpu

VEC_ROWS, POOL, DIM = 60_000, 256, 64
FACT, DIM_ROWS, GROUPS = 400_000, 20_000, 500


class _DevPtr:
    def __init__(self, ptr, rows, dim):
        self.ptr, self.shape = ptr, (rows, dim)

    def data_ptr(self):
        return self.ptr


@pytest.fixture(scope="module")
def served(backend):
    s = backend.create_stream()
    ftab = datagen.table(backend, s, datagen.filter_table_columns(seed=11), 1_000_000)
    fact_cols, dim_cols = datagen.star_columns(FACT, DIM_ROWS, GROUPS, seed=12)
    fact = datagen.table(backend, s, fact_cols, FACT)
    dim = datagen.table(backend, s, dim_cols, DIM_ROWS)
    join = backend.prepare_join_build(s, dim, 0)
    emb = datagen.table(backend, s, datagen.embedding_columns(seed=13, dim=DIM, centroids=64), VEC_ROWS + POOL)
    base = emb.column(0).values
    rows = backend.wrap([(_DevPtr(base, VEC_ROWS, DIM), None, plan.F32_VECTOR, DIM)], ["emb"])
    qv = backend.wrap([(_DevPtr(base + VEC_ROWS * DIM * 4, POOL, DIM), None, plan.F32_VECTOR, DIM)], ["emb"])
    pool = np.asarray(backend.download(s, qv).column(0).values.to_numpy(zero_copy_only=False),
                      np.float32).reshape(POOL, DIM)
    index = VectorIndex(s, rows, 0, plan.DistanceMetric.Cosine, VEC_BF16)
    s.synchronize()
    tables = serving.Tables(ftab, fact, dim, join, index, VEC_ROWS, candidates=100, k=10, keep=[emb, rows, qv])
    serving.prepare_filter_batch(backend, s, tables)  # the loop batches filter_agg requests
    yield tables, pool
    index.free()


def test_same_alone_and_in_the_mix(backend, served):
    tables, pool = served
    # arrivals far faster than service, so vector requests batch together
    reqs = serving.poisson_mix(3000.0, 150, pool, DIM_ROWS, seed=21)
    alone = copy.deepcopy(reqs)
    server = serving.Server(backend, tables)
    st = server.run(reqs)
    assert not server.errors, server.errors[:1]
    assert st["requests"] == 150
    assert max(r.batch for r in reqs if r.kind == "vector") > 1
    assert max(r.batch for r in reqs if r.kind == "filter_agg") > 1
    serving.run_one_at_a_time(backend, tables, alone)
    bad = [r.id for r, a in zip(reqs, alone) if not serving.same_result(r, a)]
    assert not bad, f"requests {bad[:10]} differ between the mix and alone"


def test_small_vector_batches_cap(backend, served):
    tables, pool = served
    reqs = serving.poisson_mix(5000.0, 80, pool, DIM_ROWS, seed=22)
    alone = copy.deepcopy(reqs)
    server = serving.Server(backend, tables, max_vector_queries=20)
    server.run(reqs)
    assert not server.errors
    assert max(r.batch for r in reqs if r.kind == "vector") <= 20  # every request holds >= 1 query
    serving.run_one_at_a_time(backend, tables, alone)
    assert all(serving.same_result(r, a) for r, a in zip(reqs, alone))


def test_aggregate_requests_match_the_oracle(backend, served, oracle):
    tables, _ = served
    s = backend.create_stream()
    host = backend.download(s, tables.filter_table)
    fact, dim = backend.download(s, tables.fact), backend.download(s, tables.dim)
    reqs = [serving.Request(0, "filter_agg", 0.0, 1), serving.Request(1, "filter_agg", 0.0, 377),
            serving.Request(2, "join_agg", 0.0, (0, 2000)), serving.Request(3, "join_agg", 0.0, (7000, 9000))]
    serving.run_one_at_a_time(backend, tables, reqs)
    for r in reqs[:2]:
        spec = serving.filter_agg_spec(r.payload)
        assert_agg_equal(r.result, oracle.aggregate(host, spec, sum_mode=1), spec)
    joined = oracle.hash_join(fact, dim, 0, 0)  # fact columns, then dim's (pk 3, grp 4)
    for r in reqs[2:]:
        js = serving.join_agg_spec(*r.payload)
        spec = AggregateSpec(4, js.aggregates, filter=js.filter)
        want = oracle.aggregate(joined, spec, sum_mode=1)
        assert r.result.num_rows > 0
        assert_agg_equal(r.result, want, spec)

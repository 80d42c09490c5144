"""The CPU oracle pinned against the reference's own known answers
(crates/oxidelake-device/src/cpu/kernels.rs tests, conformance.rs), closed
forms, brute force on tiny inputs and pyarrow compute (integer / NaN-free).
"""
import itertools
import math
import struct

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc
import pytest

from oxidelake_amd import datagen, plan
from oxidelake_amd.plan import AggregateFunction as F
from oxidelake_amd.plan import Aggregate, AggregateSpec, Lit, Predicate, col
from oxidelake_amd.plan import Comparison as Cmp

NAN = float("nan")
INF = float("inf")
I64MAX = 2**63 - 1
I64MIN = -(2**63)


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def ref_batch():
    # kernels.rs:938-973
    return pa.record_batch({
        "k": pa.array([1, 2, None, 2, 5], pa.int64()),
        "v": pa.array([0.5, None, 2.5, 3.5, -1.0], pa.float64()),
        "s": pa.array(["a", "b", "c", None, "e"]),
    })


# ------------------------------------------------------------------- filter
def test_filter_drops_null_predicates_and_projects(oracle):
    # kernels.rs:976-988
    pred = Predicate.and_(Predicate.compare(0, Cmp.GtEq, Lit.int64(2)),
                          Predicate.compare(1, Cmp.Lt, Lit.float64(4.0)))
    mask = oracle.predicate_mask(ref_batch(), pred)
    assert list(np.nonzero(mask)[0]) == [3, 4]
    out = oracle.filter_project(ref_batch().select([0, 1]), pred, [0])
    assert out.column(0).to_pylist() == [2, 5]


def test_projections_see_only_the_kept_rows(oracle):
    # a projection above a filter (DataFusion's ProjectionExec over
    # FilterExec): a zero divisor in a dropped row is no error, in a kept
    # row it is
    b = pa.record_batch({"x": pa.array([4, 5, 6], pa.int64()), "y": pa.array([2, 0, 3], pa.int64())})
    keep = Predicate.compare(1, Cmp.NotEq, Lit.int64(0))
    out = oracle.filter_project(b, keep, [plan.div(col(0), col(1))])
    assert out.column(0).to_pylist() == [2, 2]
    with pytest.raises(ZeroDivisionError):
        oracle.filter_project(b, None, [plan.div(col(0), col(1))])


def test_float_comparisons_order_as_datafusion_does(oracle):
    # kernels.rs:990-1021
    vals = [NAN, -NAN, INF, -INF, 0.0, -0.0, 4.0]
    b = pa.record_batch({"v": pa.array(vals, pa.float64())})

    def sel(op, x):
        return list(np.nonzero(oracle.predicate_mask(b, Predicate.compare(0, op, Lit.float64(x))))[0])

    assert sel(Cmp.Gt, 4.0) == [0, 2]
    assert sel(Cmp.Lt, -INF) == [1]
    assert sel(Cmp.Eq, NAN) == [0]
    assert sel(Cmp.Eq, 0.0) == [4, 5]
    assert sel(Cmp.Eq, -0.0) == [4, 5]
    assert sel(Cmp.Lt, 0.0) == [1, 3]
    assert sel(Cmp.GtEq, -0.0) == [0, 2, 4, 5, 6]


SPECIALS = [NAN, -NAN, INF, -INF, 0.0, -0.0, -1.5, 2.0, 4.0, 4.5, 5e-324, -5e-324]


def _total_cmp(a: float, b: float) -> int:
    # f64::total_cmp written from the bit patterns
    # as unsigned keys: negatives flip every bit (below 2^63), positives set the sign bit
    ka, kb = bits(a), bits(b)
    ka = (~ka) & 0xFFFFFFFFFFFFFFFF if ka >> 63 else ka | (1 << 63)
    kb = (~kb) & 0xFFFFFFFFFFFFFFFF if kb >> 63 else kb | (1 << 63)
    return (ka > kb) - (ka < kb)


def _selects(v, x, op):
    # conformance.rs:1460-1495 `selects`: -0.0 and +0.0 as one, then total_cmp
    z = lambda t: 0.0 if t == 0.0 else t  # noqa: E731
    o = _total_cmp(z(v), z(x))
    return {Cmp.Eq: o == 0, Cmp.Lt: o < 0, Cmp.LtEq: o <= 0, Cmp.Gt: o > 0,
            Cmp.GtEq: o >= 0, Cmp.NotEq: o != 0}[op]


def test_special_float_predicates_match_hand_written_order(oracle):
    # conformance.rs:1392-1458: 12 values x the 12 predicates' literals
    rng = np.random.default_rng(0x5bec1a1f)
    n = 400
    v = [SPECIALS[i] for i in rng.integers(0, len(SPECIALS), n)]
    valid = rng.random(n) >= 0.1
    b = pa.record_batch({"v": pa.array(v, pa.float64(), mask=~valid)})
    for op, x in [(Cmp.Gt, 4.0), (Cmp.GtEq, 4.5), (Cmp.Lt, 0.0), (Cmp.Gt, -0.0), (Cmp.LtEq, -0.0),
                  (Cmp.Eq, 0.0), (Cmp.Eq, -0.0), (Cmp.GtEq, NAN), (Cmp.Eq, NAN), (Cmp.Gt, INF),
                  (Cmp.Lt, -INF), (Cmp.NotEq, 2.0)]:
        got = oracle.predicate_mask(b, Predicate.compare(0, op, Lit.float64(x)))
        want = np.array([valid[i] and _selects(v[i], x, op) for i in range(n)])
        assert (got == want).all(), (op, x)


def test_kleene_logic_brute_force(oracle):
    # every combination of (true, false, NULL) for two leaves under AND / OR
    k = [1, 0, None, 1, 0, None, 1, 0, None]
    m = [1, 1, 1, 0, 0, 0, None, None, None]
    b = pa.record_batch({"k": pa.array(k, pa.int64()), "m": pa.array(m, pa.int64())})
    lk = Predicate.compare(0, Cmp.Eq, Lit.int64(1))
    lm = Predicate.compare(1, Cmp.Eq, Lit.int64(1))
    for node, fn in [(plan.and_(lk, lm), lambda a, c: a and c), (plan.or_(lk, lm), lambda a, c: a or c)]:
        got = oracle.predicate_mask(b, node)
        want = []
        for a, c in zip(k, m):
            a3 = None if a is None else a == 1
            c3 = None if c is None else c == 1
            if fn is not None and node.kind == plan.X_AND:
                r = False if (a3 is False or c3 is False) else (None if None in (a3, c3) else True)
            else:
                r = True if (a3 is True or c3 is True) else (None if None in (a3, c3) else False)
            want.append(r is True)
        assert list(got) == want


def test_extended_grammar_matches_pyarrow(oracle):
    # integer and NaN-free cases only: pyarrow's float order is IEEE, not DataFusion's
    rng = np.random.default_rng(9)
    n = 5000
    a = rng.integers(0, 1 << 40, n).astype(np.int64)
    b_ = rng.integers(0, 1000, n).astype(np.int32)
    c = rng.normal(size=n).astype(np.float32)
    cv = rng.random(n) >= 0.05
    d = np.exp(rng.normal(size=n))
    batch = pa.record_batch({"a": pa.array(a), "b": pa.array(b_), "c": pa.array(c, mask=~cv), "d": pa.array(d)})
    pred, proj = datagen.filter_query(50)
    got = oracle.filter_project(batch, pred, proj, ["a", "b2", "cd"])
    A, B, Cc, D = batch.columns
    keep = pc.or_kleene(pc.and_kleene(pc.less(B, 50), pc.greater(pc.cast(Cc, pa.float64()), 0.5)),
                        pc.equal(pc.subtract(A, pc.multiply(pc.divide(A, 7), 7)), 3))
    keep = pc.fill_null(keep, False)
    want_a = pc.filter(A, keep)
    want_b = pc.filter(pc.add(pc.multiply(pc.cast(B, pa.int64()), 2), 1), keep)
    want_cd = pc.filter(pc.multiply(pc.cast(Cc, pa.float64()), D), keep)
    assert got.column(0).equals(want_a)
    assert got.column(1).equals(want_b)
    assert got.column(2).equals(want_cd)
    assert got.column(1).type == pa.int64() and got.column(2).type == pa.float64()


def test_integer_arithmetic_wraps_and_truncates(oracle):
    b = pa.record_batch({"x": pa.array([I64MAX, -7, 7, -7, I64MIN], pa.int64()),
                         "y": pa.array([1, 2, -2, -2, -1], pa.int64())})
    add = oracle.eval_expr(b, plan.add(col(0), col(1))).to_pylist()
    assert add[0] == I64MIN  # wraps
    q = oracle.eval_expr(b.slice(1, 3), plan.div(col(0), col(1))).to_pylist()
    r = oracle.eval_expr(b.slice(1, 3), plan.mod(col(0), col(1))).to_pylist()
    assert q == [-3, -3, 3] and r == [-1, 1, -1]  # truncation toward zero
    assert oracle.eval_expr(b.slice(4, 1), plan.mod(col(0), col(1))).to_pylist() == [0]
    with pytest.raises(ZeroDivisionError):
        oracle.eval_expr(b.slice(4, 1), plan.div(col(0), col(1)))
    i32 = pa.record_batch({"j": pa.array([2**31 - 1, -(2**31)], pa.int32())})
    assert oracle.eval_expr(i32, plan.add(col(0), Lit.int32(1))).to_pylist() == [-(2**31), -(2**31) + 1]
    assert oracle.eval_expr(i32, plan.add(col(0), Lit.int64(1))).to_pylist() == [2**31, -(2**31) + 1]


# ---------------------------------------------------------------- aggregate
def test_aggregate_groups_nulls_last_and_skips_null_values(oracle):
    # kernels.rs:1123-1158
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)),
                             Aggregate(F.Min, col(1)), Aggregate(F.Max, col(0))])
    out = oracle.aggregate(ref_batch().select([0, 1]), spec)
    assert out.column(0).to_pylist() == [1, 2, 5, None]
    assert out.column(1).to_pylist() == [0.5, 3.5, -1.0, 2.5]
    assert out.column(2).to_pylist() == [1, 1, 1, 1]
    assert out.schema.field(1).name == "SUM(v)" and out.schema.field(4).name == "MAX(k)"
    assert out.column(4).to_pylist() == [1, 2, 5, None]


EDGE_ROWS = [
    (1, I64MAX - 1, 1.0), (2, 5, NAN), (None, I64MAX, -0.0), (1, I64MAX - 1, NAN),
    (3, None, None), (4, I64MIN, -NAN), (None, I64MAX, 0.0), (1, -3, 3.0),
    (4, 7, -INF), (2, -1, 0.25),
]


def edge_batch(rows):
    return pa.record_batch({
        "k": pa.array([r[0] for r in rows], pa.int64()),
        "i": pa.array([r[1] for r in rows], pa.int64()),
        "f": pa.array([r[2] for r in rows], pa.float64()),
    })


EDGE_SPEC = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)),
                              Aggregate(F.Min, col(1)), Aggregate(F.Max, col(1)),
                              Aggregate(F.Sum, col(2)), Aggregate(F.Count, col(2)),
                              Aggregate(F.Min, col(2)), Aggregate(F.Max, col(2))])


def fbits(arr):
    return [None if v is None else bits(v) for v in arr.to_pylist()]


def test_batch_by_batch_and_merged_partials_equal_one_pass(oracle):
    # kernels.rs:1209-1288
    whole = oracle.aggregate(edge_batch(EDGE_ROWS), EDGE_SPEC)
    assert whole.column(1).to_pylist() == [-7, 4, None, I64MIN + 7, -2]
    assert fbits(whole.column(7)) == [bits(1.0), bits(0.25), None, bits(-NAN), bits(-0.0)]
    assert fbits(whole.column(8)) == [bits(NAN), bits(NAN), None, bits(-INF), bits(0.0)]
    for split in range(1, len(EDGE_ROWS)):
        a, b = edge_batch(EDGE_ROWS[:split]), edge_batch(EDGE_ROWS[split:])
        streamed = oracle.GroupedAggregator(a, EDGE_SPEC)
        streamed.update(a)
        streamed.update(b)
        merged = oracle.GroupedAggregator(a, EDGE_SPEC)
        merged.merge(oracle.aggregate(b, EDGE_SPEC))
        merged.merge(oracle.aggregate(a, EDGE_SPEC))
        for got in (streamed.finish(), merged.finish()):
            for c in range(whole.num_columns):
                if whole.column(c).type == pa.float64():
                    assert fbits(got.column(c)) == fbits(whole.column(c)), (split, c)
                else:
                    assert got.column(c).to_pylist() == whole.column(c).to_pylist(), (split, c)


def test_sum_overflow_wraps_like_datafusion(oracle):
    # conformance.rs:306-356: "1|-4|2", "2|5|1", "3|MAX|2", "4|MAX-1|3", "NULL|-2|2"
    rows = [(1, I64MAX - 1), (2, 5), (1, I64MAX - 1), (3, I64MIN), (4, I64MAX), (3, -1),
            (None, I64MAX), (4, 1), (None, I64MAX), (4, -2)]
    b = pa.record_batch({"k": pa.array([r[0] for r in rows], pa.int64()),
                         "x": pa.array([r[1] for r in rows], pa.int64())})
    out = oracle.aggregate(b, AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1))]))
    rendered = ["|".join("NULL" if v is None else str(v) for v in row)
                for row in zip(*[c.to_pylist() for c in out.columns])]
    assert rendered == ["1|-4|2", "2|5|1", f"3|{I64MAX}|2", f"4|{I64MAX - 1}|3", "NULL|-2|2"]


def test_aggregate_over_empty_batch_has_no_groups(oracle):
    # kernels.rs:1304-1315
    empty = ref_batch().select([0, 1]).slice(0, 0)
    out = oracle.aggregate(empty, AggregateSpec(0, [Aggregate(F.Count, col(1))]))
    assert out.num_rows == 0 and out.num_columns == 2


def test_total_order_keys_round_trip_and_order(oracle):
    # kernels.rs:1318-1358
    vals = [struct.unpack("<d", struct.pack("<Q", b))[0] for b in (
        0xFFFFFFFFFFFFFFFF, bits(-NAN), bits(-INF), bits(-1.7976931348623157e308), bits(-1.0),
        0x8000000000000001, bits(-0.0), bits(0.0), 1, bits(2.2250738585072014e-308), bits(1.0),
        bits(1.7976931348623157e308), bits(INF), 0x7FF0000000000001, bits(NAN), 0x7FFFFFFFFFFFFFFF)]
    for a in vals:
        assert bits(oracle.from_total_order_key(oracle.total_order_key(a))) == bits(a)
        for b in vals:
            ka, kb = oracle.total_order_key(a), oracle.total_order_key(b)
            assert (ka > kb) - (ka < kb) == _total_cmp(a, b)
    assert bits(oracle.from_total_order_key(I64MAX)) == 0x7FFFFFFFFFFFFFFF
    assert bits(oracle.from_total_order_key(I64MIN)) == 0xFFFFFFFFFFFFFFFF


def test_aggregate_matches_pyarrow_group_by(oracle):
    rng = np.random.default_rng(4)
    n = 20000
    k = rng.integers(0, 300, n)
    kv = rng.random(n) >= 0.02
    x = rng.integers(-1000, 1000, n)
    xv = rng.random(n) >= 0.05
    b = pa.record_batch({"k": pa.array(k, pa.int64(), mask=~kv), "x": pa.array(x, pa.int64(), mask=~xv)})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1)), Aggregate(F.Count, col(1)), Aggregate(F.Min, col(1)),
                             Aggregate(F.Max, col(1)), Aggregate(F.CountStar)])
    out = oracle.aggregate(b, spec)
    t = pa.table({"k": b.column(0), "x": b.column(1)})
    want = t.group_by("k").aggregate([("x", "sum"), ("x", "count"), ("x", "min"), ("x", "max"),
                                      ([], "count_all")]).sort_by([("k", "ascending")])
    wk = want.column("k").to_pylist()
    gk = out.column(0).to_pylist()
    assert gk[:-1] == wk[:-1] and gk[-1] is None
    for gc, wc in [(1, "x_sum"), (2, "x_count"), (3, "x_min"), (4, "x_max"), (5, "count_all")]:
        assert out.column(gc).to_pylist() == want.column(wc).to_pylist(), wc


def test_compensated_sum_is_exact_where_sequential_drifts(oracle):
    # closed form: 1e16 + many 1.0s; the sequential fold loses every 1.0
    x = np.array([1e16] + [1.0] * 1000 + [-1e16], dtype=np.float64)
    b = pa.record_batch({"k": pa.array(np.zeros(len(x), np.int64)), "x": pa.array(x)})
    spec = AggregateSpec(0, [Aggregate(F.Sum, col(1))])
    assert oracle.aggregate(b, spec, sum_mode=0).column(1).to_pylist() == [0.0]
    assert oracle.aggregate(b, spec, sum_mode=1).column(1).to_pylist() == [1000.0]
    assert math.fsum(x) == 1000.0


# --------------------------------------------------------------------- join
def test_join_skips_null_keys_and_multiplies_matches(oracle):
    # kernels.rs:1033-1058
    right = pa.record_batch({"rk": pa.array([2, 2, None, 9], pa.int64()),
                             "name": pa.array(["two-a", "two-b", "null", "nine"])})
    out = oracle.hash_join(ref_batch(), right, 0, 0)
    assert out.num_columns == 5 and out.num_rows == 4
    assert sorted(out.column(4).to_pylist()) == ["two-a", "two-a", "two-b", "two-b"]


def test_join_matches_gather_in_ranges_equal_one_gather(oracle):
    # kernels.rs:1063-1121
    right = pa.record_batch({"rk": pa.array([2, 2, None, 9, 2, 5], pa.int64()),
                             "name": pa.array(["a", "b", "null", "nine", "c", "five"])})
    left = ref_batch()
    table = oracle.JoinTable(right, 0)
    m = oracle.JoinMatches(left, 0, table)
    assert len(m) == 7
    whole = oracle.hash_join_gather(left, right, m, 0, 7)
    proj = [0, 1, 3, 4]
    for chunk in range(1, 9):
        parts, first = [], 0
        while first < len(m):
            rows = min(chunk, len(m) - first)
            parts.append(oracle.hash_join_gather(left, right, m, first, rows, proj))
            first += rows
        joined = pa.Table.from_batches(parts).combine_chunks()
        assert joined.equals(pa.Table.from_batches([whole.select(proj)]).combine_chunks()), chunk
    with pytest.raises(IndexError):
        m.pairs(7, 1)


def test_join_equals_nested_loop(oracle):
    rng = np.random.default_rng(12)
    for n_l, n_r in [(0, 5), (7, 0), (64, 33), (300, 200)]:
        lk = rng.integers(0, 20, n_l)
        lv = rng.random(n_l) >= 0.1
        rk = rng.integers(0, 20, n_r)
        rv = rng.random(n_r) >= 0.1
        left = pa.record_batch({"k": pa.array(lk, pa.int64(), mask=~lv), "i": pa.array(np.arange(n_l))})
        right = pa.record_batch({"k": pa.array(rk, pa.int64(), mask=~rv), "j": pa.array(np.arange(n_r))})
        out = oracle.hash_join(left, right, 0, 0)
        got = list(zip(out.column(1).to_pylist(), out.column(3).to_pylist()))
        want = [(i, j) for i in range(n_l) for j in range(n_r) if lv[i] and rv[j] and lk[i] == rk[j]]
        assert got == want  # probe order, then build order


# ------------------------------------------------------------------- vector
def test_vector_distances(oracle):
    # kernels.rs:1360-1396
    child = pa.array([1.0, 0.0, 0.0, 1.0, 3.0, 4.0], pa.float32())
    vecs = pa.FixedSizeListArray.from_arrays(child, 2, mask=pa.array([False, True, False]))
    b = pa.record_batch({"vec": vecs})
    out = oracle.vector_distance(b, 0, [1.0, 0.0], plan.DistanceMetric.L2, "d")
    assert out.column(1).to_pylist() == [0.0, None, np.float32(math.sqrt(20)).item()]
    out = oracle.vector_distance(b, 0, [1.0, 0.0], plan.DistanceMetric.Cosine, "d")
    d = out.column(1).to_pylist()
    assert abs(d[0]) < 1e-6 and abs(d[2] - (1 - 3 / 5)) < 1e-6
    with pytest.raises(ValueError):
        oracle.vector_distance(b, 0, [1.0], plan.DistanceMetric.L2, "d")
    assert math.isnan(oracle.cosine(np.array([0.0, 0.0]), np.array([1.0, 0.0])))


def test_vector_distance_closed_forms(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.normal(size=37).astype(np.float32)
        q = rng.normal(size=37).astype(np.float32)
        l2 = math.sqrt(sum((float(x) - float(y)) ** 2 for x, y in zip(a, q)))
        assert abs(oracle.l2(a, q) - l2) <= 1e-5 * l2
        cs = 1 - float(a.astype(np.float64) @ q) / (np.linalg.norm(a.astype(np.float64)) * np.linalg.norm(q.astype(np.float64)))
        assert abs(oracle.cosine(a, q) - cs) <= 1e-5
    # a parallel vector has cosine distance 0; an antiparallel one 2
    a = np.array([1.0, 2.0, 2.0], np.float32)
    assert abs(oracle.cosine(a, 2 * a)) < 1e-6 and abs(oracle.cosine(a, -a) - 2) < 1e-6


def test_topk64_equals_sorting(oracle):
    rng = np.random.default_rng(2)
    v = rng.normal(size=(500, 16)).astype(np.float32)
    valid = (rng.random(500) > 0.1).astype(np.uint8)
    q = rng.normal(size=16).astype(np.float32)
    rows, dist = oracle.topk64(v, valid, q, plan.DistanceMetric.L2, 10)
    d = np.sqrt(((v.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1))
    d[valid == 0] = np.inf
    order = np.lexsort((np.arange(500), d))[:10]
    assert list(rows) == list(order)
    assert np.allclose(dist, d[order], rtol=1e-12)


def test_splitmix64_known_values():
    # the SplitMix64 reference stream (datagen.rs:89-95) from seed 0
    s, out = 0, []
    for _ in range(3):
        s, z = datagen.splitmix64(s)
        out.append(z)
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_global_aggregate_is_one_row(oracle):
    b = pa.record_batch({"x": pa.array([], pa.int64())})
    spec = AggregateSpec(-1, [Aggregate(F.Sum, col(0)), Aggregate(F.CountStar)])
    out = oracle.aggregate(b, spec)
    assert out.num_columns == 2 and [c.to_pylist() for c in out.columns] == [[None], [0]]
    b = pa.record_batch({"x": pa.array([3, None, 4], pa.int64())})
    out = oracle.aggregate(b, spec)
    assert [c.to_pylist() for c in out.columns] == [[7], [3]]

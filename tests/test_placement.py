"""The placement rule (oxidelake_amd/placement.py) on the CPU.

The v1 grammar is checked against the reference's own placement tests
(crates/oxidelake-planner/tests/placement.rs), over the same tables (t, r,
e: :27-100): the rewritten node's EXPLAIN line with its target tag, the
stock node gone, the same rows as the plain plan (`assert_same_results`,
:276-300), and the placement notes of nodes left on the CPU (:719-788). The
`rocm` target's wider grammar is checked on the configs' own query shapes:
the filter + aggregate of config 1 with its Int32 key, config 2's OR, `%`
and arithmetic projections, config 3's aggregate over a join. Lowered nodes
run through the operator over the oracle-backed fake device; the plain
plans run on pyarrow (tests/plan_exec.py).
"""
import numpy as np
import pyarrow as pa

from oxidelake_amd import placement as P
from oxidelake_amd.placement import (AggExpr, AggregateExec, Call, FilterExec, HashJoinExec, ProjectionExec,
                                     Scan, binary, col, lit)
from tests import plan_exec
from tests.cpu_path import FakeBackend, OracleCpu

T = Scan("t", [("k", P.INT64), ("v", P.FLOAT64), ("s", P.UTF8)])
R = Scan("r", [("rk", P.INT64), ("name", P.UTF8)])
E = Scan("e", [("id", P.INT64), ("emb", P.vector(2))])


def tables():
    # placement.rs:27-100 (t's two batches concatenated)
    t = pa.table({"k": pa.array([1, 2, None, 2, 5, 2, 7], pa.int64()),
                  "v": pa.array([0.5, 1.5, 2.5, None, 3.5, 4.5, -1.0], pa.float64()),
                  "s": pa.array([f"s{i}" for i in (0, 1, 2, 3, 0, 1, 2)])})
    r = pa.table({"rk": pa.array([2, 2, None, 7], pa.int64()), "name": ["two-a", "two-b", "null", "seven"]})
    emb = pa.FixedSizeListArray.from_arrays(pa.array([1.0, 0.0, 0.0, 3.0, 0.0, 0.0, 2.0, 2.0], pa.float32()), 2,
                                            mask=pa.array([False, False, True, False]))
    e = pa.table({"id": pa.array([0, 1, 2, 3], pa.int64()), "emb": emb})
    return {"t": t, "r": r, "e": e}


def lower(node, target):
    rule = P.HardwarePlacementRule(target)
    return rule.optimize(node), rule.notes


def assert_same_results(oracle, build, target, tag, data=None):
    """placement.rs:276-300: the lowered plan shows `tag` (or no Gpu node
    when tag is None) and returns the plain plan's rows."""
    data = data or tables()
    lowered, notes = lower(build(), target)
    text = P.explain(lowered)
    if tag is None:
        assert "Gpu" not in text, text
    else:
        assert tag in text, f"expected {tag} in plan:\n{text}"
    got = plan_exec.run(lowered, data, FakeBackend(oracle), OracleCpu(oracle))
    want = plan_exec.run(build(), data)
    assert plan_exec.same_rows(got, want), f"{text}\n{got}\n{want}"
    return lowered, text, notes


# ------------------------------------------------------- the v1 grammar
def test_filter_is_rewritten_and_fused_with_projection(oracle):
    # placement.rs:302-320
    q = lambda: FilterExec(T, binary("AND", binary(">=", col(T.fields, "k"), lit(2)),  # noqa: E731
                                     binary("<", col(T.fields, "v"), lit(4.0))), [2, 0])
    _, text, _ = assert_same_results(oracle, q, "cuda", "GpuFilterExec[cuda]: k >= 2 AND v < 4.0, projection=[s, k]")
    assert "FilterExec:" not in text.replace("GpuFilterExec", "")
    # literal on the left, flipped comparison
    q2 = lambda: FilterExec(T, binary("<", lit(2), col(T.fields, "k")), [0])  # noqa: E731
    assert_same_results(oracle, q2, "metal", "GpuFilterExec[metal]: k > 2")


def test_aggregate_is_rewritten_with_planned_names(oracle):
    # placement.rs:324-343
    k, v = col(T.fields, "k"), col(T.fields, "v")
    q = lambda: AggregateExec(T, [(k, "k")], [AggExpr("sum", [v], "sum(t.v)"),  # noqa: E731
                                              AggExpr("count", [v], "count(t.v)"),
                                              AggExpr("min", [v], "min(t.v)"), AggExpr("max", [k], "max(t.k)")])
    lowered, text, _ = assert_same_results(
        oracle, q, "cuda", "GpuAggregateExec[cuda]: group_by=[k], aggr=[SUM(v), COUNT(v), MIN(v), MAX(k)]")
    assert [n for n, _ in lowered.schema()] == ["k", "sum(t.v)", "count(t.v)", "min(t.v)", "max(t.k)"]
    assert "AggregateExec: mode" not in text


def _join(projection=None, join_type="Inner", filt=None):
    # SELECT t.k, t.s, r.name FROM t JOIN r ON t.k = r.rk: DataFusion builds r, probes t
    return HashJoinExec(R, T, [(col(R.fields, "rk"), col(T.fields, "k"))], join_type, filt,
                        projection if projection is not None else [2, 4, 1])


def test_join_is_rewritten(oracle):
    # placement.rs:345-353
    lowered, _, _ = assert_same_results(oracle, _join, "cuda", "GpuHashJoinExec[cuda]: join_type=Inner, on=[")
    assert [n for n, _ in lowered.schema()] == ["k", "s", "name"]
    assert lowered.line() == "GpuHashJoinExec[cuda]: join_type=Inner, on=[k = rk], projection=[k, s, name]"


def test_vector_distance_projection_is_rewritten(oracle):
    # placement.rs:639-661
    ide, emb = col(E.fields, "id"), col(E.fields, "emb")
    q = lambda: ProjectionExec(E, [(ide, "id"), (emb, "emb"),  # noqa: E731
                                   (Call("l2_distance", (emb, P.Literal([1.0, 0.0], "Vector"))), "d")])
    _, text, _ = assert_same_results(oracle, q, "cuda", "GpuVectorDistanceExec[cuda]: l2(emb) AS d, dim=2")
    assert "ProjectionExec" not in text
    q2 = lambda: ProjectionExec(E, [(ide, "id"), (emb, "emb"),  # noqa: E731
                                    (Call("cosine_distance", (P.Literal([0.0, 1.0], "Vector"), emb)), "c")])
    assert_same_results(oracle, q2, "metal", "GpuVectorDistanceExec[metal]: cosine(emb) AS c, dim=2")


def test_an_ineligible_shape_records_why_it_stayed_on_the_cpu():
    # placement.rs:719-760
    k, v, s = col(T.fields, "k"), col(T.fields, "v"), col(T.fields, "s")
    _, notes = lower(AggregateExec(T, [(k, "k")], [AggExpr("avg", [v], "avg(t.v)")]), "cuda")
    assert any(n.startswith("AggregateExec:") and "avg" in n and "sum, count, min, max" in n for n in notes), notes
    _, notes = lower(AggregateExec(T, [(s, "s")], [AggExpr("sum", [v], "sum(t.v)")]), "cuda")
    assert any("grouping key `s`" in n and "Int64" in n for n in notes), notes
    _, notes = lower(FilterExec(T, binary("OR", binary(">=", k, lit(2)), binary("<", v, lit(1.0))), [0]), "cuda")
    assert any(n.startswith("FilterExec:") for n in notes), notes
    _, notes = lower(FilterExec(T, binary("=", s, lit("s1")), [0]), "cuda")
    assert any(n.startswith("FilterExec:") and "Int64 or Float64 literal" in n for n in notes), notes


def test_a_fully_lowered_plan_records_no_notes():
    # placement.rs:765-771
    k = col(T.fields, "k")
    plan = AggregateExec(FilterExec(T, binary(">=", k, lit(2)), [0, 1]), [(k, "k")],
                         [AggExpr("sum", [col(T.fields, "v")], "sum(t.v)")])
    lowered, notes = lower(plan, "cuda")
    assert notes == [] and "Gpu" in P.explain(lowered)


def test_the_not_equal_note_names_only_comparisons_that_exist():
    # placement.rs:775-785
    _, notes = lower(FilterExec(T, binary("!=", col(T.fields, "k"), lit(2)), [0]), "cuda")
    assert notes == ["FilterExec: `!=` is not one of the comparisons the device kernels implement "
                     "(=, <, <=, >, >=)"]


def test_a_join_left_on_the_cpu_records_why():
    # placement.rs:791-830
    _, notes = lower(_join(join_type="Left"), "cuda")
    assert notes == ["HashJoinExec: a `Left` join has no device kernel; the device join is inner only"]
    _, notes = lower(_join(filt=binary("<", col(T.fields, "v"), lit(1.0))), "cuda")
    assert notes and "a condition besides its key equality" in notes[0]
    bad = HashJoinExec(R, T, [(col(R.fields, "name"), col(T.fields, "s"))], projection=[0])
    _, notes = lower(bad, "cuda")
    assert notes == ["HashJoinExec: the join key `name` is Utf8; the device join takes Int64 keys"]


def test_count_star_over_a_filter_keeps_only_the_row_count(oracle):
    # rule.rs:150-158: an empty projection keeps the predicate's first column
    # on the device and drops it above
    q = lambda: AggregateExec(FilterExec(T, binary(">=", col(T.fields, "k"), lit(2)), []), [],  # noqa: E731
                              [AggExpr("count", [lit(1)], "count(*)")])
    lowered, notes = lower(q(), "cuda")
    assert isinstance(lowered.input, ProjectionExec) and lowered.input.exprs == []
    assert isinstance(lowered.input.input, P.GpuFilterExec) and lowered.input.input.names == ["k"]
    got = plan_exec.run(lowered, tables(), FakeBackend(oracle), OracleCpu(oracle))
    assert got.column(0).to_pylist() == [5]  # k >= 2: 2, 2, 2, 5, 7


# ----------------------------------------------- the rocm target's grammar
def test_rocm_lowers_what_v1_lowers_with_its_own_tag(oracle):
    q = lambda: FilterExec(T, binary("AND", binary(">=", col(T.fields, "k"), lit(2)),  # noqa: E731
                                     binary("<", col(T.fields, "v"), lit(4.0))), [2, 0])
    assert_same_results(oracle, q, "rocm", "GpuFilterExec[rocm]: k >= 2 AND v < 4.0, projection=[s, k]")
    assert_same_results(oracle, _join, "rocm", "GpuHashJoinExec[rocm]: join_type=Inner, on=[k = rk]")


def test_rocm_lowers_or_and_not_equal(oracle):
    k, v = col(T.fields, "k"), col(T.fields, "v")
    q = lambda: FilterExec(T, binary("OR", binary(">=", k, lit(5)), binary("<", v, lit(1.0))), [0, 1])  # noqa: E731
    _, _, notes = assert_same_results(oracle, q, "rocm", "GpuFilterExec[rocm]: k >= 5 OR v < 1.0, projection=[k, v]")
    assert notes == []
    q2 = lambda: FilterExec(T, binary("!=", k, lit(2)), [0])  # noqa: E731
    assert_same_results(oracle, q2, "rocm", "GpuFilterExec[rocm]: k != 2, projection=[k]")


K1 = Scan("t1", [("id", P.INT64), ("k", P.INT32), ("v", P.FLOAT64)])


def config1(target_table="t1"):
    # SELECT k, SUM(v) FROM t WHERE k >= 2 AND v < 4.0 GROUP BY k (k Int32:
    # DataFusion coerces the literal 2 to the column's type)
    k, v = col(K1.fields, "k"), col(K1.fields, "v")
    f = FilterExec(K1, binary("AND", binary(">=", k, lit(2, P.INT32)), binary("<", v, lit(4.0))), [1, 2])
    return AggregateExec(f, [(P.Column(0, "k"), "k")], [AggExpr("sum", [P.Column(1, "v")], "sum(t.v)")])


def config1_tables(n=2000, seed=3):
    rng = np.random.default_rng(seed)
    return {"t1": pa.table({"id": pa.array(np.arange(n), pa.int64()),
                            "k": pa.array(rng.integers(0, 16, n).astype(np.int32), mask=rng.random(n) < 0.05),
                            "v": pa.array(rng.uniform(0, 8, n), mask=rng.random(n) < 0.05)})}


def test_config1_is_one_fused_aggregate_on_rocm_and_stays_on_the_cpu_under_v1(oracle):
    lowered, _, notes = assert_same_results(
        oracle, config1, "rocm", "GpuAggregateExec[rocm]: group_by=[k], aggr=[SUM(v)], filter=[k >= 2 AND v < 4.0]",
        config1_tables())
    assert notes == [] and isinstance(lowered, P.GpuAggregateExec) and isinstance(lowered.input, Scan)
    assert lowered.spec.group_by == 1 and lowered.spec.filter is not None
    _, notes = lower(config1(), "cuda")
    assert any("does not compare a column against an Int64 or Float64 literal" in n for n in notes), notes


C2 = Scan("t2", [("a", P.INT64), ("b", P.INT32), ("c", P.FLOAT32), ("d", P.FLOAT64)])


def config2(threshold=50):
    # SELECT a, b*2+1, c*d FROM t WHERE b < 50 AND c > 0.5 OR a % 7 = 3
    a, b, c, d = (col(C2.fields, x) for x in "abcd")
    pred = binary("OR", binary("AND", binary("<", b, lit(threshold, P.INT32)), binary(">", c, lit(0.5, P.FLOAT32))),
                  binary("=", binary("%", a, lit(7)), lit(3)))
    f = FilterExec(C2, pred)
    return ProjectionExec(f, [(P.Column(0, "a"), "a"),
                              (binary("+", binary("*", P.Column(1, "b"), lit(2, P.INT32)), lit(1, P.INT32)), "b2"),
                              (binary("*", P.Column(2, "c"), P.Column(3, "d")), "cd")])


def config2_tables(n=3000, seed=5):
    rng = np.random.default_rng(seed)
    return {"t2": pa.table({"a": pa.array(rng.integers(0, 2**40, n), pa.int64()),
                            "b": pa.array(rng.integers(0, 1000, n).astype(np.int32)),
                            "c": pa.array(rng.normal(size=n).astype(np.float32), mask=rng.random(n) < 0.05),
                            "d": pa.array(rng.lognormal(size=n))})}


def test_config2_is_one_filter_with_arithmetic_projections(oracle):
    lowered, text, notes = assert_same_results(
        oracle, config2, "rocm",
        "GpuFilterExec[rocm]: b < 50 AND c > 0.5 OR a % 7 = 3, projection=[a, b * 2 + 1 AS b2, c * d AS cd]",
        config2_tables())
    assert notes == [] and "ProjectionExec" not in text
    _, notes = lower(config2(), "cuda")
    assert notes and notes[0].startswith("FilterExec:")


FACT = Scan("fact", [("fk", P.INT64), ("qty", P.INT32), ("price", P.FLOAT64)])
DIM = Scan("dim", [("pk", P.INT64), ("grp", P.INT32)])


def config3():
    # SUM(qty), SUM(qty*price), COUNT(*) FROM fact JOIN dim ON fk = pk GROUP BY grp:
    # DataFusion builds dim (left) and probes fact (right); output dim ++ fact
    j = HashJoinExec(DIM, FACT, [(col(DIM.fields, "pk"), col(FACT.fields, "fk"))])
    qty, price, grp = P.Column(3, "qty"), P.Column(4, "price"), P.Column(1, "grp")
    return AggregateExec(j, [(grp, "grp")], [AggExpr("sum", [qty], "sum(qty)"),
                                             AggExpr("sum", [binary("*", qty, price)], "sum(qty * price)"),
                                             AggExpr("count", [lit(1)], "count(*)")])


def config3_tables(nf=5000, nd=300, groups=40, seed=9):
    rng = np.random.default_rng(seed)
    fk = np.minimum(rng.zipf(1.3, nf) - 1, nd + 20).astype(np.int64)  # some keys miss the dimension
    return {"fact": pa.table({"fk": pa.array(fk, mask=rng.random(nf) < 0.02),
                              "qty": pa.array(rng.integers(1, 101, nf).astype(np.int32)),
                              "price": pa.array(rng.uniform(1, 1000, nf))}),
            "dim": pa.table({"pk": pa.array(rng.permutation(nd).astype(np.int64)),
                             "grp": pa.array(rng.integers(0, groups, nd).astype(np.int32),
                                             mask=rng.random(nd) < 0.05)})}


def test_config3_is_one_fused_join_aggregate(oracle):
    lowered, _, notes = assert_same_results(
        oracle, config3, "rocm",
        "GpuJoinAggregateExec[rocm]: join_type=Inner, on=[fk = pk], group_by=[grp], "
        "aggr=[SUM(qty), SUM(qty * price), COUNT(*)]", config3_tables())
    assert notes == [] and isinstance(lowered, P.GpuJoinAggregateExec)
    assert (lowered.probe_key, lowered.build_key, lowered.build_group) == (0, 0, 1)
    # v1: the join lowers, the aggregate (an Int32 key, an expression) does not
    lowered, notes = lower(config3(), "cuda")
    assert isinstance(lowered, AggregateExec) and isinstance(lowered.input, P.GpuHashJoinExec)
    assert any("grouping key `grp` is Int32" in n for n in notes), notes


def test_rocm_still_records_what_it_cannot_lower():
    k, v, s = col(T.fields, "k"), col(T.fields, "v"), col(T.fields, "s")
    cases = [
        (AggregateExec(T, [(k, "k")], [AggExpr("avg", [v], "avg(t.v)")]),
         "AggregateExec: `avg` is not one of the aggregates the device kernels implement (sum, count, min, max)"),
        (AggregateExec(T, [(k, "k"), (v, "v")], [AggExpr("sum", [v], "s")]),
         "AggregateExec: the device kernel groups by at most one column; this groups by 2"),
        (AggregateExec(T, [(k, "k")], [AggExpr("sum", [v], "s", distinct=True)]),
         "AggregateExec: DISTINCT is not implemented for `sum`"),
        (AggregateExec(T, [(k, "k")], [AggExpr("sum", [v], "s", filter=binary(">", v, lit(0.0)))]),
         "AggregateExec: a FILTER clause on an aggregate has no device kernel"),
        (FilterExec(T, binary("=", s, lit("s1")), [0]),
         "FilterExec: `s` is Utf8; the device kernels take Int32, Int64, Float32 and Float64 columns"),
        (_join(join_type="Left"), "HashJoinExec: a `Left` join has no device kernel; the device join is inner only"),
    ]
    for plan, note in cases:
        _, notes = lower(plan, "rocm")
        assert notes == [note], (notes, note)

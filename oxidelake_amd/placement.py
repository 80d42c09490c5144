"""Device placement: which nodes of a physical plan run on the ROCm backend,
and what each lowered node hands the operator.

The reference does this in its planner, a DataFusion PhysicalOptimizerRule
written in Rust (`HardwarePlacementRule`, crates/oxidelake-planner/src/
rule.rs:37-1032): it walks the plan bottom-up, replaces each eligible node
with its `Gpu*Exec` tagged with the target (`GpuFilterExec[cuda]: ...`), and
records a placement note naming why a node it looked at stayed on the CPU.
Neither Rust nor DataFusion is available here, so this module restates that
rule over a small physical-plan IR of the same node shapes (FilterExec with
its projection, ProjectionExec, HashJoinExec with DataFusion's build-left /
probe-right sides, AggregateExec, scans), with two grammars:

- `v1`, the reference's own (targets "cuda" and "metal"): a predicate is
  `column <op> literal` (=, <, <=, >, >=; Int64 or Float64 literal of the
  column's type) joined by AND (rule.rs:797-861); an aggregate groups by one
  Int64 column and applies SUM / COUNT / MIN / MAX to plain Int64 or Float64
  columns (rule.rs:880-946); a join is inner, on one Int64 key pair, with no
  other condition (rule.rs:715-787); a distance projection appends one
  `l2_distance` / `cosine_distance` of a column and a literal query vector
  (rule.rs:243-289).
- `extended`, the target "rocm": the device grammar of this backend (the
  C ABI's postfix expressions): Int32 / Int64 / Float32 / Float64 columns
  and literals under + - * / %, all six comparisons (with !=), AND and OR,
  with DataFusion's numeric coercion; arithmetic projections fused into the
  filter; aggregates over expressions, COUNT(*) and one Int32 or Int64 key
  or none; and two fused shapes the configs need — an aggregate over a
  filter becomes one `GpuAggregateExec` with the predicate fused, and an
  aggregate over an inner join whose key groups come from the build side
  becomes one `GpuJoinAggregateExec` (the C ABI's `oxr_agg_update_join`).

Lowered nodes carry the backend's own plan objects (oxidelake_amd.plan), so
a plan lowered here runs through `GpuOperator` / `RocmBackend` unchanged.
Notes use the reference's wording where the reference has the case.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Union

from . import plan

# ------------------------------------------------------------------ types
INT32, INT64, FLOAT32, FLOAT64 = "Int32", "Int64", "Float32", "Float64"
UTF8, BOOLEAN = "Utf8", "Boolean"
NUMERIC = (INT32, INT64, FLOAT32, FLOAT64)
_PLAN_TYPE = {INT32: plan.INT32, INT64: plan.INT64, FLOAT32: plan.FLOAT32, FLOAT64: plan.FLOAT64}


def vector(dim: int) -> str:
    """FixedSizeList<Float32, dim>: a vector column's type."""
    return f"FixedSizeList<Float32, {dim}>"


def _coerce(a: str, b: str) -> str:
    """DataFusion's numeric coercion of two operand types."""
    if FLOAT64 in (a, b) or (FLOAT32 in (a, b) and INT64 in (a, b)):
        return FLOAT64
    if FLOAT32 in (a, b):
        return FLOAT32
    if INT64 in (a, b):
        return INT64
    return INT32


# ------------------------------------------------------------ expressions
@dataclass(frozen=True)
class Column:
    index: int
    name: str

    def __str__(self):
        return self.name


@dataclass(frozen=True)
class Literal:
    value: object
    dtype: str

    def __str__(self):
        if self.dtype in (FLOAT32, FLOAT64):
            return repr(float(self.value))
        if self.dtype == UTF8:
            return f"'{self.value}'"
        return str(self.value)


@dataclass(frozen=True)
class BinaryExpr:
    op: str  # = != < <= > >= AND OR + - * / %
    left: "Expr"
    right: "Expr"

    def __str__(self):
        return f"{_operand(self.left, self.op)} {self.op} {_operand(self.right, self.op)}"


@dataclass(frozen=True)
class Call:
    name: str
    args: tuple

    def __str__(self):
        return f"{self.name}({', '.join(str(a) for a in self.args)})"


Expr = Union[Column, Literal, BinaryExpr, Call]
COMPARISONS = {"=": plan.Comparison.Eq, "<": plan.Comparison.Lt, "<=": plan.Comparison.LtEq,
               ">": plan.Comparison.Gt, ">=": plan.Comparison.GtEq, "!=": plan.Comparison.NotEq}
V1_COMPARISONS = ("=", "<", "<=", ">", ">=")  # rule.rs:700-706
ARITHMETIC = {"+": plan.add, "-": plan.sub, "*": plan.mul, "/": plan.div, "%": plan.mod}
_FLIP = {"=": "=", "!=": "!=", "<": ">", "<=": ">=", ">": "<", ">=": "<="}
_PRECEDENCE = {"OR": 1, "AND": 2, "=": 3, "!=": 3, "<": 3, "<=": 3, ">": 3, ">=": 3,
               "+": 4, "-": 4, "*": 5, "/": 5, "%": 5}


def _operand(e: Expr, parent: str) -> str:
    if isinstance(e, BinaryExpr) and _PRECEDENCE[e.op] < _PRECEDENCE[parent]:
        return f"({e})"
    return str(e)


def col(schema, name: str) -> Column:
    """The column `name` of a schema (a list of (name, type))."""
    for i, (n, _) in enumerate(schema):
        if n == name:
            return Column(i, name)
    raise KeyError(name)


def lit(value, dtype: Optional[str] = None) -> Literal:
    if dtype is None:
        dtype = UTF8 if isinstance(value, str) else FLOAT64 if isinstance(value, float) else INT64
    return Literal(value, dtype)


def binary(op: str, left: Expr, right: Expr) -> BinaryExpr:
    return BinaryExpr(op, left, right)


# ------------------------------------------------------------ plan nodes
@dataclass
class Scan:
    name: str
    fields: list  # [(name, type)]

    def schema(self):
        return list(self.fields)

    def children(self):
        return []

    def line(self):
        return f"DataSourceExec: {self.name}, projection=[{', '.join(n for n, _ in self.fields)}]"


@dataclass
class FilterExec:
    input: object
    predicate: Expr
    projection: Optional[list] = None  # output columns (input indices); None: all

    def schema(self):
        s = self.input.schema()
        return s if self.projection is None else [s[i] for i in self.projection]

    def children(self):
        return [self.input]

    def line(self):
        proj = "" if self.projection is None else \
            f", projection=[{', '.join(self.input.schema()[i][0] for i in self.projection)}]"
        return f"FilterExec: {self.predicate}{proj}"


@dataclass
class ProjectionExec:
    input: object
    exprs: list  # [(expr, alias)]

    def schema(self):
        return [(alias, _type_of(e, self.input.schema())) for e, alias in self.exprs]

    def children(self):
        return [self.input]

    def line(self):
        return f"ProjectionExec: expr=[{', '.join(f'{e} as {a}' for e, a in self.exprs)}]"


@dataclass
class HashJoinExec:
    """DataFusion's HashJoinExec: builds on `left`, probes with `right`; output
    `left ++ right`, or `projection` (indices into that)."""
    left: object
    right: object
    on: list  # [(left key expr, right key expr)]
    join_type: str = "Inner"
    filter: Optional[Expr] = None
    projection: Optional[list] = None
    null_equals_null: bool = False

    def schema(self):
        full = self.left.schema() + self.right.schema()
        return full if self.projection is None else [full[i] for i in self.projection]

    def children(self):
        return [self.left, self.right]

    def line(self):
        on = ", ".join(f"({a}, {b})" for a, b in self.on)
        return f"HashJoinExec: mode=Partitioned, join_type={self.join_type}, on=[{on}]"


@dataclass
class AggExpr:
    func: str  # sum, count, min, max, avg, ...
    args: list  # expressions; count(*) is count of the literal 1 as DataFusion plans it
    alias: str
    distinct: bool = False
    filter: Optional[Expr] = None


@dataclass
class AggregateExec:
    input: object
    group_by: list  # [(expr, alias)]
    aggr: list  # [AggExpr]

    def schema(self):
        s = self.input.schema()
        out = [(a, _type_of(e, s)) for e, a in self.group_by]
        for x in self.aggr:
            t = INT64 if x.func == "count" else (FLOAT64 if x.func == "avg" else
                                                 _sum_type(_type_of(x.args[0], s)) if x.func == "sum"
                                                 else _type_of(x.args[0], s))
            out.append((x.alias, t))
        return out

    def children(self):
        return [self.input]

    def line(self):
        g = ", ".join(f"{e} as {a}" for e, a in self.group_by)
        return f"AggregateExec: mode=Single, gby=[{g}], aggr=[{', '.join(x.alias for x in self.aggr)}]"


def _sum_type(t: str) -> str:
    return INT64 if t in (INT32, INT64) else FLOAT64


def _type_of(e: Expr, schema) -> str:
    if isinstance(e, Column):
        return schema[e.index][1]
    if isinstance(e, Literal):
        return e.dtype
    if isinstance(e, BinaryExpr):
        if e.op in COMPARISONS or e.op in ("AND", "OR"):
            return BOOLEAN
        return _coerce(_type_of(e.left, schema), _type_of(e.right, schema))
    if isinstance(e, Call) and e.name in ("l2_distance", "cosine_distance"):
        return FLOAT32
    return "Unknown"


# ------------------------------------------------------- lowered nodes
@dataclass
class GpuFilterExec:
    input: object
    predicate: Expr  # as planned (for EXPLAIN)
    device_predicate: object  # oxidelake_amd.plan expression
    projection: list  # output expressions over the input: Column, or arithmetic (extended)
    names: list
    target: str

    def schema(self):
        s = self.input.schema()
        return [(n, _type_of(e, s)) for e, n in zip(self.projection, self.names)]

    def children(self):
        return [self.input]

    def device_projection(self):
        """The operator's projection: column indices, or plan expressions."""
        return [e.index if isinstance(e, Column) else _device_expr(e, self.input.schema())[0]
                for e in self.projection]

    def line(self):
        proj = ", ".join(n if isinstance(e, Column) and e.name == n else f"{e} AS {n}"
                         for e, n in zip(self.projection, self.names))
        return f"GpuFilterExec[{self.target}]: {self.predicate}, projection=[{proj}]"


@dataclass
class GpuHashJoinExec:
    """Streams `probe` (DataFusion's right), collects `build` (its left);
    output `probe ++ build` reordered by `projection` to DataFusion's
    order (rule.rs:291-390)."""
    probe: object
    build: object
    probe_key: int
    build_key: int
    projection: list
    target: str

    def schema(self):
        full = self.probe.schema() + self.build.schema()
        return [full[i] for i in self.projection]

    def children(self):
        return [self.probe, self.build]

    def line(self):
        on = f"{self.probe.schema()[self.probe_key][0]} = {self.build.schema()[self.build_key][0]}"
        return (f"GpuHashJoinExec[{self.target}]: join_type=Inner, on=[{on}], "
                f"projection=[{', '.join(n for n, _ in self.schema())}]")


@dataclass
class GpuAggregateExec:
    input: object
    spec: plan.AggregateSpec  # over the input's columns, the fused filter in spec.filter
    names: list  # output column names (key first when grouped)
    types: list
    target: str
    display: str

    def schema(self):
        return list(zip(self.names, self.types))

    def children(self):
        return [self.input]

    def line(self):
        return f"GpuAggregateExec[{self.target}]: {self.display}"


@dataclass
class GpuJoinAggregateExec:
    """An aggregate over an inner join, fused: the probe side streams, each
    match folds into the group named by the build side's group column
    (oxr_agg_update_join); nothing joined is materialised."""
    probe: object
    build: object
    probe_key: int
    build_key: int
    build_group: int
    spec: plan.AggregateSpec  # aggregates and filter over the probe side; group_by -1 (from the build)
    names: list
    types: list
    target: str
    display: str

    def schema(self):
        return list(zip(self.names, self.types))

    def children(self):
        return [self.probe, self.build]

    def line(self):
        return f"GpuJoinAggregateExec[{self.target}]: {self.display}"


@dataclass
class GpuVectorDistanceExec:
    input: object
    column: int
    query: list
    metric: plan.DistanceMetric
    alias: str
    target: str

    def schema(self):
        return self.input.schema() + [(self.alias, FLOAT32)]

    def children(self):
        return [self.input]

    def line(self):
        metric = "l2" if self.metric == plan.DistanceMetric.L2 else "cosine"
        return (f"GpuVectorDistanceExec[{self.target}]: {metric}({self.input.schema()[self.column][0]}) "
                f"AS {self.alias}, dim={len(self.query)}")


def explain(node, indent: int = 0) -> str:
    """The plan as indented lines, as DataFusion's `displayable(..).indent`."""
    lines = ["  " * indent + node.line()]
    for c in node.children():
        lines.append(explain(c, indent + 1))
    return "\n".join(lines)


# ------------------------------------------------- expression lowering
class _Refused(Exception):
    pass


def _device_expr(e: Expr, schema) -> tuple:
    """(plan expression, type) of an expression in the extended grammar, or
    _Refused with the reason."""
    if isinstance(e, Column):
        t = schema[e.index][1]
        if t not in NUMERIC:
            raise _Refused(f"`{e.name}` is {t}; the device kernels take Int32, Int64, Float32 and "
                           "Float64 columns")
        return plan.col(e.index), t
    if isinstance(e, Literal):
        if e.dtype == INT32:
            return plan.Lit.int32(int(e.value)), INT32
        if e.dtype == INT64:
            return plan.Lit.int64(int(e.value)), INT64
        if e.dtype == FLOAT32:
            return plan.Lit.float32(float(e.value)), FLOAT32
        if e.dtype == FLOAT64:
            return plan.Lit.float64(float(e.value)), FLOAT64
        raise _Refused(f"the literal {e} is {e.dtype}; the device kernels take Int32, Int64, Float32 "
                       "and Float64 literals")
    if isinstance(e, BinaryExpr):
        if e.op in ("AND", "OR"):
            l, lt = _device_expr(e.left, schema)
            r, rt = _device_expr(e.right, schema)
            if lt != BOOLEAN or rt != BOOLEAN:
                raise _Refused(f"`{e}` joins non-boolean operands with {e.op}")
            return (plan.and_(l, r) if e.op == "AND" else plan.or_(l, r)), BOOLEAN
        l, lt = _device_expr(e.left, schema)
        r, rt = _device_expr(e.right, schema)
        if BOOLEAN in (lt, rt):
            raise _Refused(f"`{e}` applies {e.op} to a boolean operand")
        if e.op in COMPARISONS:
            return plan.cmp(l, COMPARISONS[e.op], r), BOOLEAN
        if e.op in ARITHMETIC:
            return ARITHMETIC[e.op](l, r), _coerce(lt, rt)
    raise _Refused(f"`{e}` has no device kernel (the device grammar: Int32, Int64, Float32 and Float64 "
                   "columns and literals under + - * / %, comparisons, AND, OR)")


def _v1_predicate(e: Expr, schema) -> tuple:
    """(plan expression, display) in the reference's v1 grammar
    (rule.rs:797-861), or _Refused with the reference's reason."""
    if not isinstance(e, BinaryExpr):
        raise _Refused(f"`{e}` is not a binary comparison; the v1 grammar is column <op> literal, "
                       "conjoined with AND")
    if e.op == "AND":
        l, ld = _v1_predicate(e.left, schema)
        r, rd = _v1_predicate(e.right, schema)
        return plan.and_(l, r), f"{ld} AND {rd}"
    if e.op not in V1_COMPARISONS:
        raise _Refused(f"`{e.op}` is not one of the comparisons the device kernels implement "
                       f"({', '.join(V1_COMPARISONS)})")
    v1lit = lambda x: isinstance(x, Literal) and x.dtype in (INT64, FLOAT64)  # noqa: E731
    if isinstance(e.left, Column) and v1lit(e.right):
        c, l, op = e.left, e.right, e.op
    elif v1lit(e.left) and isinstance(e.right, Column):
        c, l, op = e.right, e.left, _FLIP[e.op]
    else:
        raise _Refused(f"`{e}` does not compare a column against an Int64 or Float64 literal")
    t = schema[c.index][1]
    if t != l.dtype:
        raise _Refused(f"`{c.name}` is {t} and the literal is {l.dtype}; the kernels compare a column "
                       "against a literal of its own type")
    lit_ = plan.Lit.int64(int(l.value)) if l.dtype == INT64 else plan.Lit.float64(float(l.value))
    return plan.Predicate.compare(c.index, COMPARISONS[op], lit_), f"{c.name} {op} {l}"


_FUNCS = {"sum": plan.AggregateFunction.Sum, "count": plan.AggregateFunction.Count,
          "min": plan.AggregateFunction.Min, "max": plan.AggregateFunction.Max}


def _remap(e: Expr, cols: list) -> Expr:
    """An expression over a node's output, re-expressed over its input
    (output column i is the input's Column cols[i])."""
    if isinstance(e, Column):
        return cols[e.index]
    if isinstance(e, BinaryExpr):
        return BinaryExpr(e.op, _remap(e.left, cols), _remap(e.right, cols))
    if isinstance(e, Call):
        return Call(e.name, tuple(_remap(a, cols) for a in e.args))
    return e


# ------------------------------------------------------------------ rule
@dataclass
class HardwarePlacementRule:
    """Lowers a plan bottom-up for `target`; `notes` holds one "Node: reason"
    per node it looked at and left on the CPU (rule.rs:37-100)."""
    target: str = "rocm"
    notes: list = field(default_factory=list)

    @property
    def grammar(self) -> str:
        return "extended" if self.target == "rocm" else "v1"

    def skip(self, node: str, why: str) -> None:
        self.notes.append(f"{node}: {why}")

    def optimize(self, node):
        kids = [self.optimize(c) for c in node.children()]
        node = _with_children(node, kids)
        lowered = self.lower(node)
        return node if lowered is None else lowered

    def lower(self, node):
        if isinstance(node, FilterExec):
            return self.lower_filter(node)
        if isinstance(node, ProjectionExec):
            return self.fuse_projection(node) or self.lower_vector_distance(node)
        if isinstance(node, HashJoinExec):
            return self.lower_join(node)
        if isinstance(node, AggregateExec):
            return self.lower_aggregate(node)
        return None

    # -- filter (rule.rs:137-235)
    def lower_filter(self, f: FilterExec):
        schema = f.input.schema()
        try:
            if self.grammar == "v1":
                dev, shown = _v1_predicate(f.predicate, schema)
            else:
                dev, t = _device_expr(f.predicate, schema)
                if t != BOOLEAN:
                    raise _Refused(f"`{f.predicate}` is not a predicate")
                shown = str(f.predicate)
        except _Refused as why:
            self.skip("FilterExec", str(why))
            return None
        proj = f.projection if f.projection is not None else list(range(len(schema)))
        count_only = not proj
        if count_only:
            # COUNT(*) reads no column: keep the first column the predicate
            # reads and drop it again above (rule.rs:150-158)
            proj = plan.columns_of(dev)[:1]
        exprs = [Column(i, schema[i][0]) for i in proj]
        g = GpuFilterExec(f.input, shown, dev, exprs, [schema[i][0] for i in proj], self.target)
        return ProjectionExec(g, []) if count_only else g

    def fuse_projection(self, p: ProjectionExec):
        """A projection over a lowered filter folds into it: plain columns
        without renames (rule.rs:211-235), and in the extended grammar
        arithmetic expressions and renames too."""
        g = p.input
        if not isinstance(g, GpuFilterExec) or not p.exprs:
            return None
        exprs, names = [], []
        for e, alias in p.exprs:
            if isinstance(e, Column):
                if self.grammar == "v1" and g.names[e.index] != alias:
                    return None
                inner = g.projection[e.index]
            else:
                if self.grammar == "v1":
                    return None
                if not all(isinstance(x, Column) for x in g.projection):
                    return None
                inner = _remap(e, g.projection)
                try:
                    _, t = _device_expr(inner, g.input.schema())
                except _Refused as why:
                    self.skip("ProjectionExec", str(why))
                    return None
                if t == BOOLEAN:
                    return None
            exprs.append(inner)
            names.append(alias)
        return GpuFilterExec(g.input, g.predicate, g.device_predicate, exprs, names, self.target)

    # -- vector distance (rule.rs:243-289)
    def lower_vector_distance(self, p: ProjectionExec):
        schema = p.input.schema()
        if len(p.exprs) != len(schema) + 1:
            return None
        for pos, (e, alias) in enumerate(p.exprs[:-1]):
            if not isinstance(e, Column) or e.index != pos or schema[pos][0] != alias:
                return None
        call, alias = p.exprs[-1]
        if not isinstance(call, Call) or call.name not in ("l2_distance", "cosine_distance") or len(call.args) != 2:
            return None
        a, b = call.args
        if isinstance(a, Column) and isinstance(b, Literal):
            c, q = a, b
        elif isinstance(b, Column) and isinstance(a, Literal):
            c, q = b, a
        else:
            return None
        if schema[c.index][1] != vector(len(q.value)):
            self.skip("ProjectionExec", f"the distance call is not representable: `{c.name}` is "
                                        f"{schema[c.index][1]} and the query has {len(q.value)} values")
            return None
        metric = plan.DistanceMetric.L2 if call.name == "l2_distance" else plan.DistanceMetric.Cosine
        return GpuVectorDistanceExec(p.input, c.index, list(q.value), metric, alias, self.target)

    # -- join (rule.rs:291-390, 715-787)
    def _join_refusal(self, j: HashJoinExec) -> Optional[str]:
        if j.join_type != "Inner":
            return f"a `{j.join_type}` join has no device kernel; the device join is inner only"
        if j.filter is not None:
            return "the join has a condition besides its key equality, which the device join cannot evaluate"
        if j.null_equals_null:
            return ("the join matches NULL keys to each other (IS NOT DISTINCT FROM); the device join "
                    "never matches a NULL key")
        if len(j.on) != 1:
            return f"the join has {len(j.on)} key pairs; the device join takes exactly one"
        for e, side in zip(j.on[0], (j.left, j.right)):
            if not isinstance(e, Column):
                return f"the join key `{e}` is an expression; the device join takes plain columns"
            t = side.schema()[e.index][1]
            if t != INT64:
                return f"the join key `{e.name}` is {t}; the device join takes Int64 keys"
        return None

    def lower_join(self, j: HashJoinExec):
        why = self._join_refusal(j)
        if why:
            self.skip("HashJoinExec", why)
            return None
        build, probe = j.left, j.right
        bl, pl = len(build.schema()), len(probe.schema())
        df = j.projection if j.projection is not None else list(range(bl + pl))
        # DataFusion's build ++ probe order, re-expressed over probe ++ build
        ours = [pl + i if i < bl else i - bl for i in df]
        if not ours:
            self.skip("HashJoinExec", "the join projects no column")
            return None
        return GpuHashJoinExec(probe, build, j.on[0][1].index, j.on[0][0].index, ours, self.target)

    # -- aggregate (rule.rs:491-555, 880-946)
    def lower_aggregate(self, a: AggregateExec):
        try:
            if self.grammar == "v1":
                return self._aggregate_v1(a)
            return self._aggregate_extended(a)
        except _Refused as why:
            self.skip("AggregateExec", str(why))
            return None

    def _aggregate_v1(self, a: AggregateExec):
        schema = a.input.schema()
        if len(a.group_by) != 1:
            raise _Refused(f"the device kernel groups by exactly one column; this groups by {len(a.group_by)}")
        key = a.group_by[0][0]
        if not isinstance(key, Column):
            raise _Refused("the grouping key is an expression, not a plain column")
        if schema[key.index][1] != INT64:
            raise _Refused(f"the grouping key `{key.name}` is {schema[key.index][1]}; the device kernel "
                           "groups on Int64")
        if any(x.filter is not None for x in a.aggr):
            raise _Refused("a FILTER clause on an aggregate has no device kernel")
        aggs, shown = [], []
        for x in a.aggr:
            self._check_function(x)
            if len(x.args) != 1:
                raise _Refused(f"`{x.func}` takes {len(x.args)} arguments; the device kernels take one")
            c = x.args[0]
            if not isinstance(c, Column):
                raise _Refused(f"`{x.func}` aggregates an expression, not a plain column")
            if schema[c.index][1] not in (INT64, FLOAT64):
                raise _Refused(f"`{c.name}` is {schema[c.index][1]}; the device kernels aggregate Int64 and "
                               "Float64")
            aggs.append(plan.Aggregate(_FUNCS[x.func], plan.col(c.index), x.alias))
            shown.append(f"{x.func.upper()}({c.name})")
        spec = plan.AggregateSpec(key.index, aggs)
        out = a.schema()
        return GpuAggregateExec(a.input, spec, [n for n, _ in out], [t for _, t in out], self.target,
                                f"group_by=[{key.name}], aggr=[{', '.join(shown)}]")

    def _check_function(self, x: AggExpr) -> None:
        if x.func not in _FUNCS:
            raise _Refused(f"`{x.func}` is not one of the aggregates the device kernels implement "
                           "(sum, count, min, max)")
        if x.distinct:
            raise _Refused(f"DISTINCT is not implemented for `{x.func}`")

    def _extended_aggregates(self, a: AggregateExec, schema, cols=None):
        """The aggregates over `schema` (through `cols` when the aggregate's
        input is a node that was folded away), and their display."""
        if any(x.filter is not None for x in a.aggr):
            raise _Refused("a FILTER clause on an aggregate has no device kernel")
        aggs, shown = [], []
        for x in a.aggr:
            self._check_function(x)
            star = x.func == "count" and (not x.args or (len(x.args) == 1 and isinstance(x.args[0], Literal)))
            if star:
                aggs.append(plan.Aggregate(plan.AggregateFunction.CountStar, None, x.alias))
                shown.append("COUNT(*)")
                continue
            if len(x.args) != 1:
                raise _Refused(f"`{x.func}` takes {len(x.args)} arguments; the device kernels take one")
            e = x.args[0] if cols is None else _remap(x.args[0], cols)
            dev, t = _device_expr(e, schema)
            if t == BOOLEAN:
                raise _Refused(f"`{x.func}` aggregates a predicate")
            aggs.append(plan.Aggregate(_FUNCS[x.func], dev, x.alias))
            shown.append(f"{x.func.upper()}({e if cols is None else x.args[0]})")
        return aggs, shown

    def _extended_key(self, a: AggregateExec, schema):
        if len(a.group_by) > 1:
            raise _Refused(f"the device kernel groups by at most one column; this groups by {len(a.group_by)}")
        if not a.group_by:
            return None
        key = a.group_by[0][0]
        if not isinstance(key, Column):
            raise _Refused("the grouping key is an expression, not a plain column")
        t = schema[key.index][1]
        if t not in (INT32, INT64):
            raise _Refused(f"the grouping key `{key.name}` is {t}; the device kernel groups on Int32 and Int64")
        return key

    def _aggregate_extended(self, a: AggregateExec):
        out = a.schema()
        names, types = [n for n, _ in out], [t for _, t in out]
        inp = a.input
        if isinstance(inp, GpuHashJoinExec):
            fused = self._fuse_join_aggregate(a, inp, names, types)
            if fused is not None:
                return fused
        if isinstance(inp, GpuFilterExec) and all(isinstance(e, Column) for e in inp.projection):
            # aggregate over a filter: one kernel, the predicate fused
            base = inp.input.schema()
            cols = inp.projection
            key = self._extended_key(a, inp.schema())
            aggs, shown = self._extended_aggregates(a, base, cols)
            k = -1 if key is None else cols[key.index].index
            spec = plan.AggregateSpec(k, aggs, filter=inp.device_predicate)
            kname = "" if key is None else cols[key.index].name
            return GpuAggregateExec(inp.input, spec, names, types, self.target,
                                    f"group_by=[{kname}], aggr=[{', '.join(shown)}], filter=[{inp.predicate}]")
        schema = inp.schema()
        key = self._extended_key(a, schema)
        aggs, shown = self._extended_aggregates(a, schema)
        spec = plan.AggregateSpec(-1 if key is None else key.index, aggs)
        kname = "" if key is None else key.name
        return GpuAggregateExec(inp, spec, names, types, self.target,
                                f"group_by=[{kname}], aggr=[{', '.join(shown)}]")

    def _fuse_join_aggregate(self, a: AggregateExec, j: GpuHashJoinExec, names, types):
        """An aggregate over a lowered join whose key is a build-side Int32 /
        Int64 column and whose inputs are probe-side columns; None (and the
        join stays as it is) for any other shape."""
        if len(a.group_by) != 1 or not isinstance(a.group_by[0][0], Column):
            return None
        pl = len(j.probe.schema())
        full = j.probe.schema() + j.build.schema()
        out_cols = [Column(i, full[i][0]) for i in j.projection]  # join output -> probe ++ build
        key = out_cols[a.group_by[0][0].index]
        if key.index < pl or full[key.index][1] not in (INT32, INT64):
            return None
        used = [c for x in a.aggr for arg in x.args for c in _columns(_remap(arg, out_cols))]
        if any(c.index >= pl for c in used):
            return None
        aggs, shown = self._extended_aggregates(a, j.probe.schema(), out_cols)
        spec = plan.AggregateSpec(-1, aggs)
        on = f"{j.probe.schema()[j.probe_key][0]} = {j.build.schema()[j.build_key][0]}"
        return GpuJoinAggregateExec(j.probe, j.build, j.probe_key, j.build_key, key.index - pl, spec, names,
                                    types, self.target,
                                    f"join_type=Inner, on=[{on}], group_by=[{key.name}], aggr=[{', '.join(shown)}]")


def _columns(e: Expr) -> list:
    if isinstance(e, Column):
        return [e]
    if isinstance(e, BinaryExpr):
        return _columns(e.left) + _columns(e.right)
    if isinstance(e, Call):
        return [c for x in e.args for c in _columns(x)]
    return []


def _with_children(node, kids):
    if isinstance(node, (FilterExec, ProjectionExec, AggregateExec)):
        node.input = kids[0]
    elif isinstance(node, HashJoinExec):
        node.left, node.right = kids
    return node

"""Runs a plan of oxidelake_amd.placement's IR, for the placement tests.

Stock nodes (the CPU operators DataFusion would run) are evaluated here with
pyarrow compute and plain Python — an implementation independent of the
oracle and of the device. Lowered nodes run through the operator
(`oxidelake_amd.operator.GpuOperator`) over whatever backend the test hands
it: the oracle-backed fake on the CPU, `RocmBackend` on a GPU. A fused join
+ aggregate runs on a `RocmBackend` through `AggState.update_join`; over the
fake, which has no fused kernel, as the operator's join followed by its
aggregation (the same lowering decisions, the fake's arithmetic).
"""
from __future__ import annotations

import math

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc

from oxidelake_amd import placement as P
from oxidelake_amd.operator import GpuOperator, JoinBuild
from oxidelake_amd.plan import AggregateSpec

_ARROW = {P.INT32: pa.int32(), P.INT64: pa.int64(), P.FLOAT32: pa.float32(), P.FLOAT64: pa.float64(),
          P.UTF8: pa.string()}
_CMP = {"=": pc.equal, "!=": pc.not_equal, "<": pc.less, "<=": pc.less_equal, ">": pc.greater,
        ">=": pc.greater_equal}
COUNT_ONLY = "__rows"  # the single column of an empty projection, which keeps only its row count


def _cast(a, t):
    return a if a.type == t else pc.cast(a, t)


def _arith(op, a, b):
    ta, tb = a.type, b.type
    flt = pa.types.is_floating(ta) or pa.types.is_floating(tb)
    t = pa.float64() if (flt and (ta == pa.float64() or tb == pa.float64() or pa.types.is_integer(ta)
                                  and ta.bit_width == 64 or pa.types.is_integer(tb) and tb.bit_width == 64)) \
        else pa.float32() if flt else pa.int64() if pa.int64() in (ta, tb) else pa.int32()
    a, b = _cast(a, t), _cast(b, t)
    if op == "+":
        return pc.add(a, b)
    if op == "-":
        return pc.subtract(a, b)
    if op == "*":
        return pc.multiply(a, b)
    # / and % row by row: truncating integer division, C fmod for floats
    va, vb = a.to_pylist(), b.to_pylist()
    out = []
    for x, y in zip(va, vb):
        if x is None or y is None:
            out.append(None)
        elif op == "/":
            out.append(x / y if flt else int(math.copysign(abs(x) // abs(y), x * y)))
        else:
            out.append(math.fmod(x, y) if flt else int(math.copysign(abs(x) % abs(y), x)))
    return pa.array(out, t)


def evaluate(e, batch: pa.Table):
    n = batch.num_rows
    if isinstance(e, P.Column):
        return batch.column(e.index).combine_chunks()
    if isinstance(e, P.Literal):
        return pa.array([e.value] * n, _ARROW[e.dtype])
    if isinstance(e, P.BinaryExpr):
        a, b = evaluate(e.left, batch), evaluate(e.right, batch)
        if e.op == "AND":
            return pc.and_kleene(a, b)
        if e.op == "OR":
            return pc.or_kleene(a, b)
        if e.op in _CMP:
            return _CMP[e.op](a, b)
        return _arith(e.op, a, b)
    if isinstance(e, P.Call) and e.name in ("l2_distance", "cosine_distance"):
        c = next(x for x in e.args if isinstance(x, P.Column))
        q = np.asarray(next(x for x in e.args if isinstance(x, P.Literal)).value, np.float64)
        out = []
        for row in batch.column(c.index).to_pylist():
            if row is None:
                out.append(None)
                continue
            x = np.asarray(row, np.float64)
            if e.name == "l2_distance":
                out.append(float(np.sqrt(((x - q) ** 2).sum())))
            else:
                den = np.linalg.norm(x) * np.linalg.norm(q)
                out.append(None if den == 0 else float(1 - (x @ q) / den))
        return pa.array(out, pa.float32())
    raise NotImplementedError(str(e))


def _aggregate(func, values, distinct):
    vals = [v for v in values if v is not None]
    if distinct:
        vals = list(dict.fromkeys(vals))
    if func == "count":
        return len(vals)
    if not vals:
        return None
    if func == "sum":
        return sum(vals)
    if func == "min":
        return min(vals)
    if func == "max":
        return max(vals)
    if func == "avg":
        return sum(vals) / len(vals)
    raise NotImplementedError(func)


def _named(batch, names) -> pa.Table:
    t = pa.Table.from_batches([batch]) if isinstance(batch, pa.RecordBatch) else batch
    return t.rename_columns(list(names))


def run(node, tables: dict, backend=None, cpu=None) -> pa.Table:
    """The rows `node` produces over `tables` (name -> pa.Table)."""
    op = None if backend is None else GpuOperator(backend, cpu)

    def go(n) -> pa.Table:
        if isinstance(n, P.Scan):
            return tables[n.name]
        if isinstance(n, P.FilterExec):
            t = go(n.input)
            t = t.filter(evaluate(n.predicate, t), null_selection_behavior="drop")
            return t if n.projection is None else t.select(n.projection)
        if isinstance(n, P.ProjectionExec):
            t = go(n.input)
            if not n.exprs:
                return pa.table({COUNT_ONLY: pa.nulls(t.num_rows, pa.int8())})
            return pa.table({a: evaluate(e, t) for e, a in n.exprs})
        if isinstance(n, P.HashJoinExec):
            left, right = go(n.left), go(n.right)
            lk, rk = n.on[0][0].index, n.on[0][1].index
            index = {}
            for i, k in enumerate(left.column(lk).to_pylist()):
                if k is not None:
                    index.setdefault(k, []).append(i)
            li, ri = [], []
            for j, k in enumerate(right.column(rk).to_pylist()):
                for i in index.get(k, []):
                    li.append(i)
                    ri.append(j)
            cols = [c.take(pa.array(li, pa.int64())) for c in left.columns] + \
                   [c.take(pa.array(ri, pa.int64())) for c in right.columns]
            names = left.schema.names + right.schema.names
            out = pa.table(dict(zip([f"{i}:{x}" for i, x in enumerate(names)], cols)))
            if n.projection is not None:
                out = out.select(n.projection)
            return out.rename_columns([x for x, _ in n.schema()])
        if isinstance(n, P.AggregateExec):
            t = go(n.input)
            keys = [evaluate(e, t).to_pylist() for e, _ in n.group_by]
            groups = {}
            for r in range(t.num_rows):
                groups.setdefault(tuple(k[r] for k in keys), []).append(r)
            if not n.group_by and not groups:
                groups[()] = []
            args = [[evaluate(a, t).to_pylist() for a in x.args] for x in n.aggr]
            rows = []
            for g, members in groups.items():
                row = list(g)
                for x, av in zip(n.aggr, args):
                    if x.func == "count" and (not x.args or isinstance(x.args[0], P.Literal)):
                        row.append(len(members))
                    else:
                        row.append(_aggregate(x.func, [av[0][r] for r in members], x.distinct))
                rows.append(row)
            schema = n.schema()
            return pa.table({name: pa.array([r[i] for r in rows], _ARROW.get(t_, None))
                             for i, (name, t_) in enumerate(schema)})
        # lowered nodes
        if isinstance(n, P.GpuFilterExec):
            b = go(n.input).combine_chunks().to_batches()
            b = b[0] if b else pa.RecordBatch.from_pylist([], schema=go(n.input).schema)
            out = op.filter_project(b, n.device_predicate, n.device_projection(), n.names)
            return _named(out, n.names)
        if isinstance(n, P.GpuHashJoinExec):
            probe = go(n.probe).combine_chunks().to_batches()[0]
            build = JoinBuild(go(n.build).combine_chunks().to_batches()[0], n.build_key)
            pieces = op.join_all(probe, build, n.probe_key, n.projection)
            t = pa.Table.from_batches(pieces) if pieces else None
            return _named(t, [x for x, _ in n.schema()]) if t is not None else \
                pa.table({x: pa.array([], _ARROW.get(y)) for x, y in n.schema()})
        if isinstance(n, P.GpuAggregateExec):
            b = go(n.input).combine_chunks().to_batches()[0]
            return _named(op.aggregate(b, n.spec), n.names)
        if isinstance(n, P.GpuVectorDistanceExec):
            b = go(n.input).combine_chunks().to_batches()[0]
            return _named(op.vector_distance(b, n.column, n.query, n.metric, n.alias),
                          [x for x, _ in n.schema()])
        if isinstance(n, P.GpuJoinAggregateExec):
            probe = go(n.probe).combine_chunks().to_batches()[0]
            build_b = go(n.build).combine_chunks().to_batches()[0]
            return _named(_join_aggregate(op, backend, n, probe, build_b), n.names)
        raise NotImplementedError(type(n).__name__)

    return go(node)


def _join_aggregate(op, backend, n, probe, build_b):
    from oxidelake_amd.backend import AggState, RocmBackend

    if isinstance(backend, RocmBackend):
        s = backend.create_stream()
        dp, db = backend.upload(s, probe), backend.upload(s, build_b)
        table = backend.prepare_join_build(s, db, n.build_key)
        st = AggState(backend, s, dp, n.spec)
        st.update_join(s, table, db, n.build_group, dp, n.probe_key)
        return backend.download(s, st.finish(s))
    pl = probe.num_columns
    pieces = op.join_all(probe, JoinBuild(build_b, n.build_key), n.probe_key,
                         list(range(pl)) + [pl + n.build_group])
    joined = pa.Table.from_batches(pieces).combine_chunks().to_batches()[0]
    return op.aggregate(joined, AggregateSpec(pl, n.spec.aggregates))


def same_rows(a: pa.Table, b: pa.Table, rel: float = 1e-12) -> bool:
    """Equal as multisets of rows; floats within `rel` relative."""
    if a.num_rows != b.num_rows or a.num_columns != b.num_columns:
        return False

    def key(row):
        return tuple((0, "") if v is None else (1, round(v, 6) if isinstance(v, float) else v) for v in row)

    ra = sorted(zip(*[c.to_pylist() for c in a.columns]), key=key)
    rb = sorted(zip(*[c.to_pylist() for c in b.columns]), key=key)
    for x, y in zip(ra, rb):
        for u, v in zip(x, y):
            if isinstance(u, float) and isinstance(v, float):
                if not (u == v or abs(u - v) <= rel * max(abs(u), abs(v)) or (math.isnan(u) and math.isnan(v))):
                    return False
            elif u != v:
                return False
    return True

"""Plan vocabulary: predicates, scalar expressions and aggregate specs.

Mirrors crates/oxidelake-core/src/params.rs:12-149 (`Comparison`, `Literal`,
`Predicate`, `AggregateFunction`, `AggregateSpec`, `DistanceMetric`) and
extends it with what the benchmark queries need: `Or`, Int32/Float32
literals, arithmetic (+ - * / %) and expression operands, COUNT(*).

Expressions encode to the C ABI's postfix `oxr_expr_node` arrays
(include/oxide_rocm.h); the encoding is the same for the device library and
for the CPU oracle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Sequence, Union

# ---------------------------------------------------------------- ABI enums
INT32, INT64, FLOAT32, FLOAT64, F32_VECTOR = 1, 2, 3, 4, 5

X_COLUMN, X_LITERAL, X_ADD, X_SUB, X_MUL, X_DIV, X_MOD, X_CMP, X_AND, X_OR = range(1, 11)


class Comparison(IntEnum):
    """params.rs:12-24 (plus NotEq, which DataFusion has and v1 lowers nowhere)."""

    Eq = 1
    Lt = 2
    LtEq = 3
    Gt = 4
    GtEq = 5
    NotEq = 6

    @property
    def symbol(self) -> str:
        return {1: "=", 2: "<", 3: "<=", 4: ">", 5: ">=", 6: "!="}[int(self)]


class AggregateFunction(IntEnum):
    """params.rs:117-127 plus COUNT(*)."""

    Sum = 1
    Count = 2
    Min = 3
    Max = 4
    CountStar = 5

    @property
    def sql(self) -> str:
        return {1: "SUM", 2: "COUNT", 3: "MIN", 4: "MAX", 5: "COUNT"}[int(self)]


class DistanceMetric(IntEnum):
    """params.rs:151-157."""

    L2 = 1
    Cosine = 2


class ExprNode(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("type", C.c_int32),
        ("column", C.c_int32),
        ("op", C.c_int32),
        ("lit_i", C.c_int64),
        ("lit_f", C.c_double),
    ]


class Expr(C.Structure):
    _fields_ = [("nodes", C.POINTER(ExprNode)), ("count", C.c_int32)]


# ------------------------------------------------------------- expressions
@dataclass(frozen=True)
class Col:
    index: int


@dataclass(frozen=True)
class Lit:
    value: Union[int, float]
    type: int = INT64

    @staticmethod
    def int64(v: int) -> "Lit":
        return Lit(int(v), INT64)

    @staticmethod
    def int32(v: int) -> "Lit":
        return Lit(int(v), INT32)

    @staticmethod
    def float64(v: float) -> "Lit":
        return Lit(float(v), FLOAT64)

    @staticmethod
    def float32(v: float) -> "Lit":
        return Lit(float(v), FLOAT32)


@dataclass(frozen=True)
class Bin:
    kind: int
    left: "ExprT"
    right: "ExprT"
    op: int = 0


ExprT = Union[Col, Lit, Bin]


def col(i: int) -> Col:
    return Col(i)


def add(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_ADD, a, b)


def sub(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_SUB, a, b)


def mul(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_MUL, a, b)


def div(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_DIV, a, b)


def mod(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_MOD, a, b)


def cmp(a: ExprT, op: Comparison, b: ExprT) -> Bin:
    return Bin(X_CMP, a, b, int(op))


def and_(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_AND, a, b)


def or_(a: ExprT, b: ExprT) -> Bin:
    return Bin(X_OR, a, b)


class Predicate:
    """Builders in the reference's own spelling (params.rs:76-87)."""

    @staticmethod
    def compare(column: int, op: Comparison, literal: Lit) -> Bin:
        return cmp(Col(column), op, literal)

    @staticmethod
    def and_(left: ExprT, right: ExprT) -> Bin:
        return and_(left, right)

    @staticmethod
    def or_(left: ExprT, right: ExprT) -> Bin:
        return or_(left, right)


def postfix(e: ExprT) -> list[ExprNode]:
    out: list[ExprNode] = []

    def walk(x: ExprT) -> None:
        if isinstance(x, Col):
            out.append(ExprNode(X_COLUMN, 0, x.index, 0, 0, 0.0))
        elif isinstance(x, Lit):
            if x.type in (INT32, INT64):
                out.append(ExprNode(X_LITERAL, x.type, 0, 0, int(x.value), 0.0))
            else:
                out.append(ExprNode(X_LITERAL, x.type, 0, 0, 0, float(x.value)))
        elif isinstance(x, Bin):
            walk(x.left)
            walk(x.right)
            out.append(ExprNode(x.kind, 0, 0, x.op, 0, 0.0))
        else:
            raise TypeError(f"not an expression: {x!r}")

    walk(e)
    return out


class Encoded:
    """An expression as a C `oxr_expr`, keeping its node array alive."""

    def __init__(self, e: ExprT | None):
        nodes = postfix(e) if e is not None else []
        self.nodes = (ExprNode * max(1, len(nodes)))(*nodes)
        self.expr = Expr(C.cast(self.nodes, C.POINTER(ExprNode)), len(nodes))

    def ref(self):
        return C.byref(self.expr)


def encode_many(exprs: Sequence[ExprT]) -> tuple[C.Array, list[Encoded]]:
    enc = [Encoded(e) for e in exprs]
    arr = (Expr * max(1, len(enc)))(*[x.expr for x in enc])
    return arr, enc


def columns_of(e: ExprT) -> list[int]:
    """Column indices referenced, in evaluation order (params.rs:90-104)."""
    if isinstance(e, Col):
        return [e.index]
    if isinstance(e, Bin):
        return columns_of(e.left) + columns_of(e.right)
    return []


def leaf_count(e: ExprT) -> int:
    """Comparison leaves (params.rs:107-113)."""
    if isinstance(e, Bin):
        if e.kind == X_CMP:
            return 1
        if e.kind in (X_AND, X_OR):
            return leaf_count(e.left) + leaf_count(e.right)
    return 0


@dataclass
class Aggregate:
    func: AggregateFunction
    input: ExprT | None = None  # None for COUNT(*)
    name: str | None = None


@dataclass
class AggregateSpec:
    """params.rs:142-149, generalised: key column (-1 = one global group),
    aggregates over expressions, an optional fused filter."""

    group_by: int
    aggregates: list[Aggregate] = field(default_factory=list)
    filter: ExprT | None = None
    groups_hint: int = 0

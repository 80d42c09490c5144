"""Seeded synthetic tables for the benchmark workloads, generated in HBM.

Each table is a list of `GenColumn` specs handed to the library's counter-
based generator (oxr_generate: SplitMix64 draws per (seed, column, row)), so
the full-size tables never cross PCIe. Tests copy them to the host
(`RocmBackend.download`) to run the CPU oracle on the same rows. The
reference's own demo generator is SplitMix64 too
(crates/oxidelake-storage/src/datagen.rs:89-95); `splitmix64` restates it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import pyarrow as pa

from . import plan

MASK64 = (1 << 64) - 1

GEN_UNIFORM_INT, GEN_UNIFORM_REAL, GEN_NORMAL, GEN_LOGNORMAL, GEN_ZIPF, GEN_SEQUENCE, \
    GEN_PERMUTATION, GEN_MIXTURE, GEN_MIXTURE_NORMALIZED = range(1, 10)


def splitmix64(state: int) -> tuple[int, int]:
    """datagen.rs:89-95: returns (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


@dataclass
class GenColumn:
    name: str
    type: int
    dist: int
    a: float = 0.0
    b: float = 0.0
    null_frac: float = 0.0
    dim: int = 0
    seed: int = 42


# ---------------------------------------------------------------- configs
def groupby_table_columns(seed: int = 42) -> list[GenColumn]:
    """1M-row table: id Int64 0..N, k Int32 U[0,16), v Float64 U[0,8),
    emb FixedSizeList<Float32, 8> i.i.d. N(0,1); no NULLs."""
    return [GenColumn("id", plan.INT64, GEN_SEQUENCE, 0, 0, seed=seed),
            GenColumn("k", plan.INT32, GEN_UNIFORM_INT, 0, 16, seed=seed),
            GenColumn("v", plan.FLOAT64, GEN_UNIFORM_REAL, 0.0, 8.0, seed=seed),
            GenColumn("emb", plan.F32_VECTOR, GEN_MIXTURE, 0, 1.0, dim=8, seed=seed)]


def groupby_query_spec() -> plan.AggregateSpec:
    """SELECT k, SUM(v) FROM t WHERE k >= 2 AND v < 4.0 GROUP BY k (k = col 1, v = col 2)."""
    C = plan.Comparison
    return plan.AggregateSpec(
        1, [plan.Aggregate(plan.AggregateFunction.Sum, plan.col(2))],
        filter=plan.and_(plan.cmp(plan.col(1), C.GtEq, plan.Lit.int64(2)),
                         plan.cmp(plan.col(2), C.Lt, plan.Lit.float64(4.0))))


def filter_table_columns(seed: int = 42, null_frac: float = 0.05) -> list[GenColumn]:
    """a Int64 U[0,2^40), b Int32 U[0,1000), c Float32 N(0,1) with 5% NULLs,
    d Float64 lognormal(0,1)."""
    return [GenColumn("a", plan.INT64, GEN_UNIFORM_INT, 0, float(1 << 40), seed=seed),
            GenColumn("b", plan.INT32, GEN_UNIFORM_INT, 0, 1000, seed=seed),
            GenColumn("c", plan.FLOAT32, GEN_NORMAL, 0.0, 1.0, null_frac=null_frac, seed=seed),
            GenColumn("d", plan.FLOAT64, GEN_LOGNORMAL, 0.0, 1.0, seed=seed)]


def filter_query(b_threshold: int = 50, c_threshold: float = 0.5, modulus: int = 7):
    """SELECT a, b*2+1, c*d FROM t WHERE b < T AND c > 0.5 OR a % 7 = 3 over
    columns (a, b, c, d) = (0, 1, 2, 3). Returns (predicate, projection)."""
    C = plan.Comparison
    pred = plan.or_(
        plan.and_(plan.cmp(plan.col(1), C.Lt, plan.Lit.int64(b_threshold)),
                  plan.cmp(plan.col(2), C.Gt, plan.Lit.float64(c_threshold))),
        plan.cmp(plan.mod(plan.col(0), plan.Lit.int64(modulus)), C.Eq, plan.Lit.int64(3)),
    )
    proj = [plan.col(0),
            plan.add(plan.mul(plan.col(1), plan.Lit.int64(2)), plan.Lit.int64(1)),
            plan.mul(plan.col(2), plan.col(3))]
    return pred, proj


def filter_selectivity(b_threshold: int, c_threshold: float = 0.5, null_frac: float = 0.05,
                        modulus: int = 7) -> float:
    """Expected selectivity: P(A) + 1/m - P(A)/m with P(A) = T/1000 * (1-nulls) * P(c > t)."""
    from math import erf, sqrt

    pc = 0.5 * (1 - erf(c_threshold / sqrt(2)))
    pa_ = (b_threshold / 1000) * (1 - null_frac) * pc
    return pa_ + 1 / modulus - pa_ / modulus


def star_columns(fact_rows: int, dim_rows: int, groups: int, seed: int = 42):
    """fact: fk Int64 Zipf(1.1) over dim_rows keys, qty Int32 U[1,100],
    price Float64 U[1,1000]; dim: pk Int64 a permutation of 0..dim_rows-1,
    grp Int32 U[0, groups)."""
    fact = [GenColumn("fk", plan.INT64, GEN_ZIPF, 1.1, float(dim_rows), seed=seed),
            GenColumn("qty", plan.INT32, GEN_UNIFORM_INT, 1, 101, seed=seed),
            GenColumn("price", plan.FLOAT64, GEN_UNIFORM_REAL, 1.0, 1000.0, seed=seed)]
    dim = [GenColumn("pk", plan.INT64, GEN_PERMUTATION, 0, 0, seed=seed + 1),
           GenColumn("grp", plan.INT32, GEN_UNIFORM_INT, 0, groups, seed=seed + 1)]
    return fact, dim


def star_query_spec() -> plan.AggregateSpec:
    """SUM(qty), SUM(qty*price), COUNT(*) over the fact columns (qty = 1,
    price = 2); the group key comes from the dimension side."""
    F = plan.AggregateFunction
    return plan.AggregateSpec(-1, [
        plan.Aggregate(F.Sum, plan.col(1), "SUM(qty)"),
        plan.Aggregate(F.Sum, plan.mul(plan.col(1), plan.col(2)), "SUM(qty*price)"),
        plan.Aggregate(F.CountStar, None, "COUNT(*)"),
    ])


def embedding_columns(seed: int = 42, dim: int = 768, centroids: int = 1024, sigma: float = 0.3):
    """Embeddings: a Gaussian mixture of N(0,1) centroids, sigma per point,
    L2-normalised."""
    return [GenColumn("emb", plan.F32_VECTOR, GEN_MIXTURE_NORMALIZED, centroids, sigma, dim=dim, seed=seed)]


def table(backend, stream, columns: list[GenColumn], rows: int):
    return backend.generate(stream, columns, rows)


# ------------------------------------------------------------ host helpers
def random_batch(rng: np.random.Generator, n: int, null_frac: float = 0.1,
                 specials: bool = True) -> pa.RecordBatch:
    """A mixed-type batch for parity tests: i Int64, j Int32, f Float32,
    g Float64 (with NaN, ±0, ±inf, subnormals when `specials`), each with
    NULLs, plus a vector column e FixedSizeList<Float32, 3>."""
    i = rng.integers(-50, 50, n).astype(np.int64)
    j = rng.integers(-20, 20, n).astype(np.int32)
    f = rng.normal(size=n).astype(np.float32)
    g = rng.normal(scale=3, size=n)
    if specials and n:
        sp = np.array([np.nan, -np.nan, np.inf, -np.inf, 0.0, -0.0, 5e-324, -5e-324, 4.0, 4.5])
        idx = rng.integers(0, n, max(1, n // 8))
        g[idx] = sp[rng.integers(0, len(sp), len(idx))]
        f[idx] = sp[rng.integers(0, len(sp), len(idx))].astype(np.float32)
    masks = [rng.random(n) < null_frac for _ in range(5)]
    e = rng.normal(size=(n, 3)).astype(np.float32)
    ev = pa.FixedSizeListArray.from_arrays(pa.array(e.reshape(-1), pa.float32()), 3)
    if n:
        ev = pa.FixedSizeListArray.from_arrays(pa.array(e.reshape(-1), pa.float32()), 3,
                                               mask=pa.array(masks[4]))
    return pa.record_batch({
        "i": pa.array(i, pa.int64(), mask=masks[0]),
        "j": pa.array(j, pa.int32(), mask=masks[1]),
        "f": pa.array(f, pa.float32(), mask=masks[2]),
        "g": pa.array(g, pa.float64(), mask=masks[3]),
        "e": ev,
    })

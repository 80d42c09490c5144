"""GPU parity: vector distance and top-k vs the CPU oracle.

Distances are FP32 sums in another order than the CPU's sequential fold, so
they are checked against an FP64 recomputation, as the reference's own
conformance suite does (crates/oxidelake-compute/tests/conformance.rs:
653-738, absolute 1e-4 at dimension 4): here |d - d64| <= 1e-4 * max(1, d64)
at every dimension. Top-k is exact up to distance ties: every row returned
is within that tolerance of the FP64 oracle's k-th distance, the rows equal
the oracle's wherever the oracle's distances are separated by more than the
tolerance, and distances come out sorted. Reduced-precision scans (BF16,
FP8 e4m3 with per-row scales) with an FP32 rerank of 100 candidates are
checked for recall@10 >= 0.99 against the FP64 oracle.
"""
import math

import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import datagen, plan
from oxidelake_amd.backend import VEC_BF16, VEC_FP8, VEC_FP32, PlanError, VectorIndex
from oxidelake_amd.plan import DistanceMetric as M

pytestmark = pytest.mark.gpu

TOL = 1e-4


def vec_batch(vecs: np.ndarray, valid=None) -> pa.RecordBatch:
    n, dim = vecs.shape
    child = pa.array(vecs.reshape(-1), pa.float32())
    mask = None if valid is None else pa.array(~valid.astype(bool))
    col = pa.FixedSizeListArray.from_arrays(child, dim, mask=mask)
    return pa.record_batch({"id": pa.array(np.arange(n, dtype=np.int64)), "v": col})


def dist64(vecs, q, metric):
    v = vecs.astype(np.float64)
    q = q.astype(np.float64)
    if metric == M.Cosine:
        na = np.linalg.norm(v, axis=1)
        nb = np.linalg.norm(q)
        with np.errstate(invalid="ignore", divide="ignore"):
            d = 1.0 - (v @ q) / (na * nb)
        d[(na == 0) | (nb == 0)] = np.nan
        return d
    return np.sqrt(((v - q) ** 2).sum(1))


def test_reference_known_answers(backend, stream, oracle):
    # kernels.rs:1360-1396: L2 = [0, NULL, sqrt(20)], cosine row 2 = 1 - 3/5
    child = pa.array([1.0, 0.0, 0.0, 1.0, 3.0, 4.0], pa.float32())
    vecs = pa.FixedSizeListArray.from_arrays(child, 2, mask=pa.array([False, True, False]))
    b = pa.record_batch({"vec": vecs})
    dev = backend.upload(stream, b)
    got = backend.download(stream, backend.vector_distance(stream, dev, 0, [1.0, 0.0], M.L2, "d"))
    assert got.schema.names == ["vec", "d"]
    d = got.column(1).to_pylist()
    assert d[0] == 0.0 and d[1] is None and abs(d[2] - math.sqrt(20)) < 1e-6
    got = backend.download(stream, backend.vector_distance(stream, dev, 0, [1.0, 0.0], M.Cosine, "d"))
    d = got.column(1).to_pylist()
    assert abs(d[0]) < 1e-6 and d[1] is None and abs(d[2] - (1 - 3 / 5)) < 1e-6
    with pytest.raises(PlanError):
        backend.vector_distance(stream, dev, 0, [1.0], M.L2, "d")
    # a zero vector has no cosine distance
    z = pa.record_batch({"vec": pa.FixedSizeListArray.from_arrays(pa.array([0.0, 0.0], pa.float32()), 2)})
    got = backend.download(stream, backend.vector_distance(stream, backend.upload(stream, z), 0, [1.0, 0.0],
                                                           M.Cosine, "d"))
    assert math.isnan(got.column(1).to_pylist()[0])


@pytest.mark.parametrize("dim", [3, 8, 129, 768])
@pytest.mark.parametrize("metric", [M.L2, M.Cosine])
def test_distance_parity(backend, stream, oracle, dim, metric):
    rng = np.random.default_rng(dim * 7 + int(metric))
    n = 5003
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    valid = rng.random(n) > 0.1
    b = vec_batch(vecs, valid)
    q = rng.normal(size=dim).astype(np.float32)
    got = backend.download(stream, backend.vector_distance(stream, backend.upload(stream, b), 1, q, metric, "dist"))
    want = oracle.vector_distance(b, 1, q, metric, "dist")
    # the input columns pass through untouched, the distance is appended
    assert got.schema.names == ["id", "v", "dist"]
    assert got.column(0).equals(b.column(0)) and got.column(1).equals(b.column(1))
    gd = got.column(2)
    assert gd.is_valid().equals(want.column(2).is_valid())
    g = gd.fill_null(0.0).to_numpy()
    d64 = dist64(vecs, q, metric)
    ok = valid
    assert np.all(np.abs(g[ok] - d64[ok]) <= TOL * np.maximum(1.0, np.abs(d64[ok])))
    # and against the reference's own FP32 fold
    w = want.column(2).fill_null(0.0).to_numpy()
    assert np.all(np.abs(g[ok] - w[ok]) <= TOL * np.maximum(1.0, np.abs(w[ok])))


def check_topk(rows, dist, vecs, valid, q, metric, k):
    d64 = dist64(vecs, q, metric)
    d64[~valid] = np.nan
    finite = ~np.isnan(d64)
    order = np.lexsort((np.arange(len(d64)), np.where(finite, d64, np.inf)))
    want = order[: min(k, int(finite.sum()))]
    got = rows[rows >= 0]
    assert len(got) == len(want)
    assert np.all(rows[len(want):] == -1)
    # sorted, exact distances within tolerance
    gd = dist[: len(want)].astype(np.float64)
    assert np.all(np.diff(gd) >= 0)
    assert np.all(np.abs(gd - d64[got]) <= TOL * np.maximum(1.0, np.abs(d64[got])))
    if len(want) == 0:
        return
    kth = d64[want[-1]]
    tol = TOL * max(1.0, abs(kth))
    # every returned row is as near as the oracle's k-th, up to ties
    assert np.all(d64[got] <= kth + tol)
    # rows clearly inside the oracle's top-k are all returned
    must = want[d64[want] < kth - 2 * tol]
    assert set(must.tolist()) <= set(got.tolist())


@pytest.mark.parametrize("n,dim", [(1, 8), (100, 8), (5000, 64), (70_001, 768)])
@pytest.mark.parametrize("metric", [M.L2, M.Cosine])
def test_exact_topk(backend, stream, n, dim, metric):
    rng = np.random.default_rng(n + dim)
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    valid = rng.random(n) > 0.05
    if n > 10:
        vecs[3] = 0.0  # zero norm: no cosine distance, never returned
    dev = backend.upload(stream, vec_batch(vecs, valid))
    for nq, k in [(1, 10), (3, 1), (40, 10)]:
        qs = rng.normal(size=(nq, dim)).astype(np.float32)
        rows, dist = backend.vector_topk(stream, dev, 1, qs, metric, k)
        vv = valid.copy()
        if metric == M.Cosine and n > 10:
            vv[3] = False
        for i in range(nq):
            check_topk(rows[i], dist[i], vecs, vv, qs[i], metric, k)


def test_many_queries_in_several_passes(backend, stream):
    # more than 256 queries: the scan runs in passes of 256
    rng = np.random.default_rng(5)
    n, dim = 20_000, 32
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    valid = np.ones(n, bool)
    dev = backend.upload(stream, vec_batch(vecs))
    qs = rng.normal(size=(300, dim)).astype(np.float32)
    ix = VectorIndex(stream, dev, 1, M.L2, VEC_FP32)
    rows, dist = ix.search(stream, qs, 5)
    for i in range(0, 300, 37):
        check_topk(rows[i], dist[i], vecs, valid, qs[i], M.L2, 5)


def mixture(backend, stream, n, dim, normalized, seed=42):
    cols = [datagen.GenColumn("v", plan.F32_VECTOR,
                              datagen.GEN_MIXTURE_NORMALIZED if normalized else datagen.GEN_MIXTURE,
                              64, 0.3, dim=dim, seed=seed)]
    dev = datagen.table(backend, stream, cols, n)
    host = backend.download(stream, dev)
    vecs = np.asarray(host.column(0).values.to_numpy(zero_copy_only=False), np.float32).reshape(n, dim)
    return dev, vecs


@pytest.mark.parametrize("precision", [VEC_BF16, VEC_FP8])
@pytest.mark.parametrize("metric", [M.L2, M.Cosine])
def test_reduced_precision_recall(backend, stream, oracle, precision, metric):
    n, dim, nq, k = 60_000, 768, 32, 10
    dev, vecs = mixture(backend, stream, n, dim, normalized=metric == M.Cosine)
    rng = np.random.default_rng(11)
    qs = vecs[rng.integers(0, n, nq)] + rng.normal(scale=0.3, size=(nq, dim)).astype(np.float32)
    if metric == M.Cosine:
        qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    qs = qs.astype(np.float32)
    ix = VectorIndex(stream, dev, 0, metric, precision)
    rows, dist = ix.search(stream, qs, k, candidates=100)
    hit = 0
    for i in range(nq):
        want, wd = oracle.topk64(vecs, None, qs[i], metric, k)
        hit += len(set(want.tolist()) & set(rows[i].tolist()))
        # reranked distances are exact FP32 distances of the rows returned
        d64 = dist64(vecs[rows[i]], qs[i], metric)
        assert np.all(np.abs(dist[i] - d64) <= TOL * np.maximum(1.0, np.abs(d64)))
    assert hit / (nq * k) >= 0.99, hit / (nq * k)


def test_index_outlives_its_batch(backend, stream):
    import gc

    rng = np.random.default_rng(9)
    vecs = rng.normal(size=(3000, 16)).astype(np.float32)
    dev = backend.upload(stream, vec_batch(vecs))
    ix = VectorIndex(stream, dev, 1, M.L2, VEC_BF16)
    del dev
    gc.collect()
    q = vecs[17] + 0.01
    rows, dist = ix.search(stream, q, 3, candidates=50)
    assert rows[0][0] == 17
    # a distance batch's input columns are views that outlive the input
    dev = backend.upload(stream, vec_batch(vecs))
    out = backend.vector_distance(stream, dev, 1, q, M.L2, "d")
    del dev
    gc.collect()
    got = backend.download(stream, out)
    assert got.column(0).to_pylist()[:3] == [0, 1, 2]


def test_empty_and_all_null(backend, stream):
    dev = backend.upload(stream, vec_batch(np.zeros((0, 4), np.float32)))
    rows, dist = backend.vector_topk(stream, dev, 1, np.ones(4, np.float32), M.L2, 3)
    assert rows.tolist() == [[-1, -1, -1]]
    vecs = np.ones((10, 4), np.float32)
    dev = backend.upload(stream, vec_batch(vecs, np.zeros(10, bool)))
    rows, dist = backend.vector_topk(stream, dev, 1, np.ones(4, np.float32), M.L2, 3)
    assert rows.tolist() == [[-1, -1, -1]]
    out = backend.download(stream, backend.vector_distance(stream, dev, 1, [1, 1, 1, 1], M.L2, "d"))
    assert out.column(2).null_count == 10


@pytest.mark.parametrize("precision", [VEC_FP32, VEC_BF16, VEC_FP8])
@pytest.mark.parametrize("overflow", [False, True])
def test_pilot_bound_matches_plain_scan(backend, stream, precision, overflow):
    # the first rows' best C bound each query's C-th best from below; the
    # other rows only append keys at or above it. Same keys, same results as
    # the plain scan, bit for bit; an overflowing list falls back to it
    from oxidelake_amd import autotune
    from oxidelake_amd._capi import lib

    n, dim = 60_000, 64
    dev, vecs = mixture(backend, stream, n, dim, normalized=True, seed=9)
    rng = np.random.default_rng(3)
    ix = VectorIndex(stream, dev, 0, M.Cosine, precision)
    for nq in (1, 40, 300):
        qs = (vecs[rng.integers(0, n, nq)] + rng.normal(scale=0.2, size=(nq, dim))).astype(np.float32)
        autotune.set_knob("vector.pilot", 0)
        plain = ix.search(stream, qs, 10, candidates=100)
        autotune.set_knob("vector.pilot", 1)
        autotune.set_knob("vector.pilot_rows", 4096)
        if overflow:
            autotune.set_knob("vector.pilot_cap", 16)
        lib.oxr_profile_enable(1)
        lib.oxr_profile_reset()
        try:
            got = ix.search(stream, qs, 10, candidates=100)
            scans = autotune.launches("vector_scan")
        finally:
            lib.oxr_profile_enable(0)
            lib.oxr_profile_reset()
            for k in ("vector.pilot", "vector.pilot_rows", "vector.pilot_cap"):
                autotune.set_knob(k, -1)
        assert np.array_equal(got[0], plain[0]) and np.array_equal(got[1], plain[1]), nq
        chunks = -(-nq // 256)
        # pilot + appending scan (+ the plain scan after an overflow) per chunk of 256 queries
        assert scans == chunks * (3 if overflow else 2), scans


@pytest.mark.parametrize("precision,knobs", [
    (VEC_FP8, {"vector.gemm": 0, "vector.wide": 1}), (VEC_BF16, {"vector.gemm": 0, "vector.wide": 3}),
    (VEC_FP32, {"vector.gemm": 0, "vector.nt8": 1}), (VEC_BF16, {"vector.gemm": 0, "vector.nt8": 1}),
    (VEC_FP8, {"vector.gemm": 1}), (VEC_BF16, {"vector.gemm": 1}), (VEC_FP32, {"vector.gemm": 1})])
def test_wide_query_blocks_match_streamed(backend, stream, oracle, precision, knobs):
    # other tile shapes for large batches against the streamed 4-tile kernel
    # in launches of 128 (vector.gemm = 0, vector.wide = 0): FP8 with 128
    # queries resident, 2 row tiles x 4 query tiles per wave; BF16 with 64
    # resident, 4 x 2 per wave, a six-deep load rotation; the 8-tile
    # streamed kernel in one launch of 256; the GEMM tile (256 rows x 256
    # queries per block, rows and queries staged through LDS). The same keys
    # (the same k groups in the same order into the same products), so the
    # same results bit for bit, and recall against the FP64 oracle
    from oxidelake_amd import autotune

    n, dim, nq, k = 40_000, 768, 200, 10
    dev, vecs = mixture(backend, stream, n, dim, normalized=True, seed=21)
    rng = np.random.default_rng(8)
    qs = vecs[rng.integers(0, n, nq)] + rng.normal(scale=0.3, size=(nq, dim)).astype(np.float32)
    qs = (qs / np.linalg.norm(qs, axis=1, keepdims=True)).astype(np.float32)
    ix = VectorIndex(stream, dev, 0, M.Cosine, precision)
    cand = {VEC_BF16: 100, VEC_FP8: 1000, VEC_FP32: k}[precision]
    autotune.set_knob("vector.pilot", 0)
    try:
        for kn, v in knobs.items():
            autotune.set_knob(kn, v)
        wide = ix.search(stream, qs, k, candidates=cand)
        for kn in knobs:
            autotune.set_knob(kn, -1)
        autotune.set_knob("vector.gemm", 0)
        autotune.set_knob("vector.wide", 0)
        streamed = ix.search(stream, qs, k, candidates=cand)
    finally:
        for kn in ("vector.pilot", "vector.wide", "vector.gemm", *knobs):
            autotune.set_knob(kn, -1)
    assert np.array_equal(wide[0], streamed[0]) and np.array_equal(wide[1], streamed[1])
    hit = 0
    for i in range(0, nq, 10):
        want, _ = oracle.topk64(vecs, None, qs[i], M.Cosine, k)
        hit += len(set(want.tolist()) & set(wide[0][i].tolist()))
    assert hit / (len(range(0, nq, 10)) * k) >= 0.99


@pytest.mark.parametrize("precision", [VEC_FP32, VEC_BF16, VEC_FP8])
@pytest.mark.parametrize("metric", [M.L2, M.Cosine])
def test_gemm_scan_matches_streamed(backend, stream, precision, metric):
    """The GEMM-tiled scan (129-256 queries) against the streamed 4-tile
    kernel, bit for bit: a ragged row count (the last block partly past the
    end), NULL and zero rows, 129 and 256 queries, in score mode (plain scan)
    and in append mode (a small pilot bounds the rest)."""
    from oxidelake_amd import autotune

    n, dim, k = 70_001, 768, 10
    rng = np.random.default_rng(31 + precision)
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    vecs[5] = 0.0
    valid = rng.random(n) > 0.03
    dev = backend.upload(stream, vec_batch(vecs, valid))
    ix = VectorIndex(stream, dev, 1, metric, precision)
    cand = {VEC_BF16: 100, VEC_FP8: 300, VEC_FP32: k}[precision]
    for nq in (129, 256):
        qs = (vecs[rng.integers(0, n, nq)] + rng.normal(scale=0.5, size=(nq, dim))).astype(np.float32)
        for pilot in (0, 1):
            got = {}
            for g in (1, 0):
                autotune.set_knob("vector.gemm", g)
                autotune.set_knob("vector.wide", 1 if g else 0)
                autotune.set_knob("vector.pilot", pilot)
                autotune.set_knob("vector.pilot_rows", 8192)
                try:
                    got[g] = ix.search(stream, qs, k, candidates=cand)
                finally:
                    for kn in ("vector.gemm", "vector.wide", "vector.pilot", "vector.pilot_rows"):
                        autotune.set_knob(kn, -1)
            assert np.array_equal(got[1][0], got[0][0]), (nq, pilot)
            assert np.array_equal(got[1][1], got[0][1], equal_nan=True), (nq, pilot)
            assert not np.isin(got[1][0], np.nonzero(~valid)[0]).any()


@pytest.mark.parametrize("precision", [VEC_FP32, VEC_BF16, VEC_FP8])
@pytest.mark.parametrize("table", ["mixture", "sparse", "ties"])
def test_sampled_selection_matches_streaming(backend, stream, precision, table):
    """Selection by a sampled bound then one sort (vector.select_sampled = 1:
    for every candidate count) against the streaming selection, bit for bit,
    on the plain scan and the
    pilot path: a mixture table; one with 97% NULL rows (the sample finds no
    bound: the streaming fallback); one of identical rows (every score ties,
    so the bound keeps more keys than the buffer holds: the fallback)."""
    from oxidelake_amd import autotune

    rng = np.random.default_rng(17 + precision)
    n, dim, k = 200_003, 64, 10
    vecs = rng.normal(size=(n, dim)).astype(np.float32)
    valid = np.ones(n, bool)
    if table == "sparse":
        valid = rng.random(n) < 0.03
    elif table == "ties":
        vecs[:] = vecs[0]
    dev = backend.upload(stream, vec_batch(vecs, valid))
    ix = VectorIndex(stream, dev, 1, M.Cosine, precision)
    cand = {VEC_BF16: 100, VEC_FP8: 1000, VEC_FP32: k}[precision]
    for nq in (1, 40, 256):
        qs = (vecs[rng.integers(0, n, nq)] + rng.normal(scale=0.3, size=(nq, dim))).astype(np.float32)
        for pilot in (0, 1):
            got = {}
            for sampled in (1, 0):
                autotune.set_knob("vector.select_sampled", sampled)
                autotune.set_knob("vector.pilot", pilot)
                autotune.set_knob("vector.pilot_rows", 16384)
                try:
                    got[sampled] = ix.search(stream, qs, k, candidates=cand)
                finally:
                    for kn in ("vector.select_sampled", "vector.pilot", "vector.pilot_rows"):
                        autotune.set_knob(kn, -1)
            assert np.array_equal(got[1][0], got[0][0]), (nq, pilot)
            assert np.array_equal(got[1][1], got[0][1], equal_nan=True), (nq, pilot)

"""A serving loop over HBM-resident tables: mixed requests, continuous
batching, admission by device-memory footprint, overlap on the device.

Three request kinds, over tables that stay resident for the server's life:
  - "vector": top-k of 1..64 queries over a vector index (BF16 scan + FP32
    rerank by default);
  - "filter_agg": COUNT(*), SUM(a), SUM(c*d) FROM t WHERE b < T over the
    200M-row filter table: every request waiting when the worker becomes
    free (up to 16) runs as one pass over the table, the run's chain
    replayed from a captured hipGraph (RocmBackend.prepare_agg_batch);
  - "join_agg": SUM(qty), SUM(qty*price), COUNT(*) FROM fact JOIN dim ON
    fk = pk WHERE fk in [lo, hi) GROUP BY grp.

Every kind has its own worker thread and HIP stream, so requests of
different kinds run at the same time on the device. Requests are admitted
continuously: a vector worker takes every vector request waiting when it
becomes free (up to 256 queries) into one scan, and the filter_agg worker
every waiting filter_agg request into one shared scan, so requests that
arrive while a batch runs join the next one instead of queueing one by one. A
memory gate admits a batch only when its estimated device footprint fits in
the budget beside everything in flight. The one-at-a-time path
(`run_one_at_a_time`) runs the same requests in arrival order, each alone,
on one stream: the plain drop-in path.

Requests run as prepared statements (one compiled kernel per query shape,
the literals passed as arguments) with compensated FP64 sums, so that a
request's result does not depend on what runs beside it.

Nothing here is a second implementation of an operator: every request runs
through RocmBackend's entry points (the library's HIP kernels).
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import pyarrow as pa

from . import autotune, plan
from .backend import AggBatch, AggState, JoinTable, RocmBackend, VectorIndex
from .plan import Aggregate, AggregateSpec
from .plan import AggregateFunction as F
from .plan import Comparison as Cmp


@dataclass
class Request:
    id: int
    kind: str                 # "vector" | "filter_agg" | "join_agg"
    arrival: float            # seconds after the run starts
    payload: object           # vector: (q, dim) float32; filter_agg: T; join_agg: (lo, hi)
    result: object = None
    start: float = 0.0        # when its batch was admitted to the device
    done: float = 0.0
    batch: int = 1            # requests in the batch that served it

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class Tables:
    """The resident state a server answers from."""
    filter_table: object      # DeviceBatch: a Int64, b Int32, c Float32 (NULLs), d Float64
    fact: object              # DeviceBatch: fk Int64, qty Int32, price Float64
    dim: object               # DeviceBatch: pk Int64, grp Int32
    join: JoinTable           # built once over dim.pk
    index: VectorIndex
    vec_rows: int             # rows the index holds
    candidates: int = 100     # rerank depth of the vector index
    vec_dim: int = 768        # the index's vector size
    k: int = 10
    keep: list = field(default_factory=list)  # anything the tables view into
    filter_batch: Optional[AggBatch] = None   # batched filter_agg requests (prepare_filter_batch)


def prepare_filter_batch(be: RocmBackend, stream, tables: Tables) -> AggBatch:
    """Prepares the filter_agg shape over the filter table for batched runs
    (the Server's filter_agg worker uses it when present)."""
    serving_knobs()
    tables.filter_batch = be.prepare_agg_batch(stream, tables.filter_table, filter_agg_spec(1))
    return tables.filter_batch


def filter_agg_spec(threshold: int) -> AggregateSpec:
    return AggregateSpec(-1, [Aggregate(F.CountStar, None, "COUNT(*)"), Aggregate(F.Sum, plan.col(0), "SUM(a)"),
                              Aggregate(F.Sum, plan.mul(plan.col(2), plan.col(3)), "SUM(c*d)")],
                         filter=plan.cmp(plan.col(1), Cmp.Lt, plan.Lit.int64(threshold)))


def join_agg_spec(lo: int, hi: int) -> AggregateSpec:
    return AggregateSpec(-1, [Aggregate(F.Sum, plan.col(1), "SUM(qty)"),
                              Aggregate(F.Sum, plan.mul(plan.col(1), plan.col(2)), "SUM(qty*price)"),
                              Aggregate(F.CountStar, None, "COUNT(*)")],
                         filter=plan.and_(plan.cmp(plan.col(0), Cmp.GtEq, plan.Lit.int64(lo)),
                                          plan.cmp(plan.col(0), Cmp.Lt, plan.Lit.int64(hi))))


def serving_knobs() -> None:
    """Prepared statements (one compiled kernel per query shape, literals as
    arguments) and compensated FP64 sums (a request's result does not depend
    on what runs beside it)."""
    autotune.set_knob("jit.params", 1)
    autotune.set_knob("agg.exact_f64", 1)


class MemoryGate:
    """Admission by footprint: acquire(n) waits until n more bytes fit under
    the budget beside everything admitted (one request always fits alone)."""

    def __init__(self, budget: int):
        self.budget = budget
        self.used = 0
        self.peak = 0
        self.cv = threading.Condition()

    def acquire(self, n: int) -> None:
        with self.cv:
            while self.used > 0 and self.used + n > self.budget:
                self.cv.wait()
            self.used += n
            self.peak = max(self.peak, self.used)

    def release(self, n: int) -> None:
        with self.cv:
            self.used -= n
            self.cv.notify_all()


def vector_scratch(rows: int, dim: int, nq: int, candidates: int) -> int:
    """Device scratch of one top-k search (oxidelake_amd/csrc/vector.hip
    `search`): queries and results, then per pass of up to 256 queries the
    prepared queries and either the pilot path (an index of at least 4 pilots
    of max(1M, 2048 C) rows: the pilot's score tile and segment candidates,
    the bests, the append list of max(64K, 4 C x pilots in the rest) keys per
    query) or the plain path (score tiles of at most 2 GiB, segment
    candidates of at least 32 C rows). A pilot whose list overflows reruns
    the plain path on top: rare, and outside this estimate."""
    C = max(candidates, 10)
    total = nq * dim * 4 + nq * 12 * 16
    pilot = max(1 << 20, 2048 * C)
    for q0 in range(0, nq, 256):
        q = min(256, nq - q0)
        total += q * dim * 4 + q * 4 + 2 * q * C * 8  # prepared queries, scales, bests

        def scan_select(n):
            tile = min(n, max(128, (2 << 30) // (4 * q)))
            seg = max(8192, 32 * C, -(-tile // max(1, 2048 // q)))
            segs = -(-n // tile) * -(-tile // seg)
            return q * tile * 4 + q * segs * C * 8

        if rows >= 4 * pilot:
            rest = rows - pilot
            cap = max(65536, 4 * C * -(-rest // pilot))
            total += scan_select(pilot) + q * C * 8 + q * cap * 8 + q * 4
        else:
            total += scan_select(rows)
    return total


def footprint(tables: Tables, kind: str, reqs: list[Request]) -> int:
    """Estimated device bytes a batch holds while it runs (scratch and
    results; the resident tables are not counted)."""
    if kind == "vector":
        nq = sum(len(r.payload) for r in reqs)
        return vector_scratch(tables.vec_rows, tables.vec_dim, nq, tables.candidates)
    if kind == "join_agg":
        # a slot-to-group entry per dimension row and one group state per row
        return tables.dim.num_rows * (8 + 64)
    return 1 << 20


def execute(be: RocmBackend, tables: Tables, stream, kind: str, reqs: list[Request], plain: bool = False) -> list:
    """Runs one batch on `stream`; returns one result per request. `plain`:
    aggregate requests through the plain create/update/finish chain, one
    by one, even where a batched form is prepared."""
    if kind == "vector":
        qs = np.concatenate([r.payload for r in reqs])
        rows, dist = tables.index.search(stream, qs, tables.k, tables.candidates)
        out, at = [], 0
        for r in reqs:
            q = len(r.payload)
            out.append((rows[at:at + q], dist[at:at + q]))
            at += q
        return out
    if kind == "filter_agg" and tables.filter_batch is not None and not plain:
        return tables.filter_batch.run(stream, [filter_agg_spec(int(r.payload)) for r in reqs])
    out = []
    for r in reqs:
        if kind == "filter_agg":
            st = AggState(be, stream, tables.filter_table, filter_agg_spec(int(r.payload)))
            st.update(stream, tables.filter_table)
        else:
            lo, hi = r.payload
            st = AggState(be, stream, tables.fact, join_agg_spec(lo, hi))
            st.update_join(stream, tables.join, tables.dim, 1, tables.fact, 0)
        out.append(be.download(stream, st.finish(stream)))
    return out


class Server:
    """Continuous batching over resident tables; see the module docstring."""

    KINDS = ("vector", "filter_agg", "join_agg")

    def __init__(self, be: RocmBackend, tables: Tables, budget: Optional[int] = None,
                 max_vector_queries: int = 256):
        self.be = be
        self.tables = tables
        # default: half of the device memory the resident tables left free
        self.gate = MemoryGate(budget if budget is not None else be.memory_info().free_bytes // 2)
        self.max_q = max_vector_queries
        self.queues = {k: queue.Queue() for k in self.KINDS}
        self.held: dict[str, Request] = {}  # a request that did not fit its kind's last batch
        self.streams = {k: be.create_stream() for k in self.KINDS}
        self.t0 = 0.0
        self.errors: list = []

    def _take(self, kind: str) -> Optional[list[Request]]:
        q = self.queues[kind]
        first = self.held.pop(kind, None)
        if first is None:
            first = q.get()
        if first is None:
            return None
        batch = [first]
        if kind == "filter_agg" and self.tables.filter_batch is not None:
            while len(batch) < self.tables.filter_batch.max_requests:  # every one waiting now
                try:
                    nxt = q.get_nowait()
                except queue.Empty:
                    break
                if nxt is None:
                    q.put(None)
                    break
                batch.append(nxt)
        if kind == "vector":
            nq = len(first.payload)
            while True:  # every vector request waiting now, up to max_q queries
                try:
                    nxt = q.get_nowait()
                except queue.Empty:
                    break
                if nxt is None:  # the end marker stays for the next take
                    q.put(None)
                    break
                if nq + len(nxt.payload) > self.max_q:
                    self.held[kind] = nxt  # heads the next batch: order within a kind is kept
                    break
                batch.append(nxt)
                nq += len(nxt.payload)
        return batch

    def _worker(self, kind: str) -> None:
        stream = self.streams[kind]
        while True:
            batch = self._take(kind)
            if batch is None:
                return
            n = footprint(self.tables, kind, batch)
            self.gate.acquire(n)
            start = time.perf_counter() - self.t0
            try:
                results = execute(self.be, self.tables, stream, kind, batch)
            except Exception as e:  # a failed request fails alone
                self.errors.append(e)
                results = [e] * len(batch)
            finally:
                self.gate.release(n)
            done = time.perf_counter() - self.t0
            for r, res in zip(batch, results):
                r.result, r.start, r.done, r.batch = res, start, done, len(batch)

    def run(self, requests: list[Request]) -> dict:
        """Feeds `requests` at their arrival times and serves them; returns
        latency and throughput statistics."""
        serving_knobs()
        workers = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in self.KINDS]
        self.t0 = time.perf_counter()
        for w in workers:
            w.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            wait = r.arrival - (time.perf_counter() - self.t0)
            if wait > 0:
                time.sleep(wait)
            self.queues[r.kind].put(r)
        for k in self.KINDS:
            self.queues[k].put(None)
        for w in workers:
            w.join()
        return stats(requests, time.perf_counter() - self.t0)


def run_one_at_a_time(be: RocmBackend, tables: Tables, requests: list[Request]) -> dict:
    """The plain path: one stream, arrival order, every request alone."""
    serving_knobs()
    stream = be.create_stream()
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        r.start = time.perf_counter() - t0
        r.result = execute(be, tables, stream, r.kind, [r], plain=True)[0]
        r.done = time.perf_counter() - t0
        r.batch = 1
    return stats(requests, time.perf_counter() - t0)


def stats(requests: list[Request], elapsed: float) -> dict:
    lat = np.array([r.latency for r in requests]) * 1000
    out = {"requests": len(requests), "elapsed_s": elapsed, "throughput_rps": len(requests) / elapsed,
           "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99))}
    for k in Server.KINDS:
        kl = np.array([r.latency for r in requests if r.kind == k]) * 1000
        if len(kl):
            out[k] = {"requests": int(len(kl)), "latency_ms_p50": float(np.percentile(kl, 50)),
                      "latency_ms_p99": float(np.percentile(kl, 99)),
                      "mean_batch": float(np.mean([r.batch for r in requests if r.kind == k]))}
    return out


def poisson_mix(rate: float, n: int, query_pool: np.ndarray, dim_rows: int, seed: int = 7,
                max_queries: int = 64) -> list[Request]:
    """n requests arriving as a Poisson process at `rate` per second: 60%
    vector (1 to max_queries queries from the pool), 30% filter_agg (T in
    [1, 500]: 0.1% to 50% of the rows), 10% join_agg (an fk range of 1-10%
    of the keys)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        u = rng.random()
        if u < 0.6:
            q = int(rng.integers(1, max_queries + 1))
            at = rng.integers(0, len(query_pool), q)
            out.append(Request(i, "vector", float(t[i]), query_pool[at]))
        elif u < 0.9:
            out.append(Request(i, "filter_agg", float(t[i]), int(rng.integers(1, 501))))
        else:
            w = int(dim_rows * rng.uniform(0.01, 0.1))
            lo = int(rng.integers(0, max(1, dim_rows - w)))
            out.append(Request(i, "join_agg", float(t[i]), (lo, lo + w)))
    return out


def same_result(a: Request, b: Request, rel: float = 1e-12, dist_tol: float = 1e-6) -> bool:
    """Whether two runs of one request agree: top-k rows equal except where
    distances tie within `dist_tol` (then the distances must agree);
    aggregate integers exact, floats within `rel` relative (a float sum's
    order of additions depends on scheduling)."""
    if isinstance(a.result, Exception) or isinstance(b.result, Exception):
        return False
    if a.kind == "vector":
        (ra, da), (rb, db) = a.result, b.result
        if ra.shape != rb.shape or not np.allclose(da, db, rtol=0, atol=dist_tol):
            return False
        for qa, qb, dq in zip(ra, rb, da):
            for i in np.nonzero(qa != qb)[0]:
                if not (np.abs(dq - dq[i]) <= dist_tol).sum() > 1:  # a row only swaps with a tie
                    return False
        return True
    ta, tb = a.result, b.result
    if ta.num_rows != tb.num_rows or ta.schema != tb.schema:
        return False
    for x, y in zip(ta.columns, tb.columns):
        if not np.array_equal(x.is_null().to_numpy(zero_copy_only=False),
                              y.is_null().to_numpy(zero_copy_only=False)):
            return False
        zero = 0.0 if pa.types.is_floating(x.type) else 0
        xv = x.fill_null(zero).to_numpy(zero_copy_only=False)
        yv = y.fill_null(zero).to_numpy(zero_copy_only=False)
        if xv.dtype.kind == "f":
            ok = np.isnan(xv) == np.isnan(yv)
            scale = np.maximum(np.abs(np.nan_to_num(xv)), 1.0)
            if not (ok.all() and np.all(np.abs(np.nan_to_num(xv) - np.nan_to_num(yv)) <= rel * scale)):
                return False
        elif not np.array_equal(xv, yv):
            return False
    return True


def describe_difference(a: Request, b: Request) -> str:
    """What differs between two runs of one request (for reports)."""
    if isinstance(a.result, Exception) or isinstance(b.result, Exception):
        return f"{a.kind}: error {a.result!r} / {b.result!r}"
    if a.kind == "vector":
        (ra, da), (rb, db) = a.result, b.result
        rows = int((ra != rb).sum())
        return (f"vector ({len(ra)} queries): {rows} rows differ, max distance difference "
                f"{float(np.abs(da - db).max()):.3g}, distances {da[ra != rb][:4].tolist()} vs "
                f"{db[ra != rb][:4].tolist()}")
    parts = []
    for name, x, y in zip(a.result.schema.names, a.result.columns, b.result.columns):
        xv = np.asarray(x.to_numpy(zero_copy_only=False), np.float64)
        yv = np.asarray(y.to_numpy(zero_copy_only=False), np.float64)
        if xv.shape != yv.shape:
            parts.append(f"{name}: {xv.shape} vs {yv.shape} rows")
            continue
        with np.errstate(all="ignore"):
            rel = np.nan_to_num(np.abs(xv - yv) / np.maximum(np.abs(yv), 1e-300))
        if rel.size and rel.max() > 0:
            i = int(rel.argmax())
            parts.append(f"{name}: max relative difference {rel[i]:.3g} ({xv[i]!r} vs {yv[i]!r})")
    return f"{a.kind} {a.payload!r}: " + ("; ".join(parts) or "same values")

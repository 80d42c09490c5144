"""Host logic of the serving loop, on the CPU: continuous batching of vector
requests (at most 256 queries, order within a kind kept), admission by
memory footprint, statistics and result comparison. Requests run through a
stand-in for the device call here; tests/test_serving_gpu.py runs them on
the device."""
import threading
import time

import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd import serving
from oxidelake_amd.serving import MemoryGate, Request


class _Mem:
    free_bytes = 1 << 30


class _Backend:
    def create_stream(self):
        return object()

    def memory_info(self):
        return _Mem()


class _Tables:
    vec_rows = 1000
    vec_dim = 8
    candidates = 10
    filter_batch = None  # filter_agg requests one by one

    class dim:
        num_rows = 10


class _BatchedTables(_Tables):
    class filter_batch:  # a prepared batch of at most 8 filter_agg requests
        max_requests = 8


def _fake_execute(log, delay=0.002):
    def run(be, tables, stream, kind, reqs, plain=False):
        log.append((kind, [r.id for r in reqs]))
        time.sleep(delay)
        return [("done", r.id) for r in reqs]

    return run


def _mix(n, rate, seed=0):
    pool = np.zeros((64, 4), np.float32)
    return serving.poisson_mix(rate, n, pool, dim_rows=1000, seed=seed)


def test_every_request_answered_once_and_vector_batches_capped(monkeypatch):
    log = []
    monkeypatch.setattr(serving, "execute", _fake_execute(log, delay=0.01))
    reqs = _mix(300, rate=20000.0)  # arrivals far faster than service: batches form
    server = serving.Server(_Backend(), _Tables(), max_vector_queries=40)
    st = server.run(reqs)
    assert st["requests"] == 300 and not server.errors
    seen = [i for _, ids in log for i in ids]
    assert sorted(seen) == list(range(300))
    by_id = {r.id: r for r in reqs}
    for kind, ids in log:
        assert all(by_id[i].kind == kind for i in ids)
        if kind == "vector":
            # capped, except a lone request larger than the cap
            assert len(ids) == 1 or sum(len(by_id[i].payload) for i in ids) <= 40
        else:
            assert len(ids) == 1  # aggregates run one request per call
    # within a kind, batches keep arrival order
    for kind in serving.Server.KINDS:
        order = [i for k, ids in log if k == kind for i in ids]
        assert order == sorted(order)
    assert any(len(ids) > 1 for k, ids in log if k == "vector")
    for r in reqs:
        assert r.result == ("done", r.id) and r.done >= r.start >= 0


def test_filter_agg_requests_batched_when_prepared(monkeypatch):
    log = []
    monkeypatch.setattr(serving, "execute", _fake_execute(log, delay=0.01))
    reqs = _mix(300, rate=20000.0, seed=5)
    server = serving.Server(_Backend(), _BatchedTables(), max_vector_queries=40)
    server.run(reqs)
    assert not server.errors
    assert sorted(i for _, ids in log for i in ids) == list(range(300))
    fa = [ids for k, ids in log if k == "filter_agg"]
    assert all(1 <= len(ids) <= 8 for ids in fa) and any(len(ids) > 1 for ids in fa)
    order = [i for ids in fa for i in ids]
    assert order == sorted(order)
    assert all(len(ids) == 1 for k, ids in log if k == "join_agg")


def test_one_at_a_time_runs_in_arrival_order(monkeypatch):
    log = []
    monkeypatch.setattr(serving, "execute", _fake_execute(log, delay=0.0))
    reqs = _mix(50, rate=5000.0, seed=3)
    st = serving.run_one_at_a_time(_Backend(), _Tables(), reqs)
    assert [ids[0] for _, ids in log] == sorted(r.id for r in reqs)
    assert all(len(ids) == 1 for _, ids in log)
    assert st["latency_ms_p99"] >= st["latency_ms_p50"] >= 0


def test_memory_gate_admits_within_budget():
    gate = MemoryGate(100)
    gate.acquire(60)
    got = []

    def second():
        gate.acquire(60)  # does not fit beside the first
        got.append(gate.used)
        gate.release(60)

    t = threading.Thread(target=second)
    t.start()
    time.sleep(0.05)
    assert got == []
    gate.release(60)
    t.join(5)
    assert got == [60] and gate.peak == 60 and gate.used == 0
    # a request larger than the whole budget still runs, alone
    gate.acquire(500)
    assert gate.used == 500
    gate.release(500)


def test_a_failed_request_fails_alone(monkeypatch):
    def run(be, tables, stream, kind, reqs):
        if kind == "join_agg":
            raise RuntimeError("boom")
        return [r.id for r in reqs]

    monkeypatch.setattr(serving, "execute", run)
    reqs = _mix(60, rate=5000.0, seed=5)
    server = serving.Server(_Backend(), _Tables())
    server.run(reqs)
    for r in reqs:
        if r.kind == "join_agg":
            assert isinstance(r.result, RuntimeError)
        else:
            assert r.result == r.id
    assert server.errors


def test_poisson_mix_shape():
    reqs = _mix(2000, rate=100.0, seed=9)
    kinds = [r.kind for r in reqs]
    assert 0.55 < kinds.count("vector") / 2000 < 0.65
    assert 0.25 < kinds.count("filter_agg") / 2000 < 0.35
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0) and 15 < arr[-1] < 25  # ~2000 / 100 s
    for r in reqs:
        if r.kind == "vector":
            assert 1 <= len(r.payload) <= 64
        elif r.kind == "filter_agg":
            assert 1 <= r.payload <= 500
        else:
            lo, hi = r.payload
            assert 0 <= lo < hi <= 1000


def _req(kind, result):
    r = Request(0, kind, 0.0, None)
    r.result = result
    return r


def test_same_result_vector_ties_and_tolerance():
    rows = np.array([[5, 7, 9]])
    d = np.array([[0.1, 0.2, 0.2]], np.float32)
    a = _req("vector", (rows, d))
    assert serving.same_result(a, _req("vector", (np.array([[5, 9, 7]]), d)))  # tie swap
    assert not serving.same_result(a, _req("vector", (np.array([[7, 5, 9]]), d)))
    assert not serving.same_result(a, _req("vector", (rows, d + 1e-3)))
    assert not serving.same_result(a, _req("vector", RuntimeError("x")))


def test_same_result_aggregates():
    t = pa.record_batch({"n": pa.array([3], pa.int64()), "s": pa.array([1.0e6], pa.float64()),
                         "m": pa.array([None], pa.int64())})
    close = pa.record_batch({"n": pa.array([3], pa.int64()), "s": pa.array([1.0e6 * (1 + 1e-14)]),
                             "m": pa.array([None], pa.int64())})
    far = pa.record_batch({"n": pa.array([3], pa.int64()), "s": pa.array([1.0e6 * (1 + 1e-9)]),
                           "m": pa.array([None], pa.int64())})
    other = pa.record_batch({"n": pa.array([4], pa.int64()), "s": pa.array([1.0e6]),
                             "m": pa.array([None], pa.int64())})
    nulls = pa.record_batch({"n": pa.array([3], pa.int64()), "s": pa.array([1.0e6]),
                             "m": pa.array([0], pa.int64())})
    a = _req("filter_agg", t)
    assert serving.same_result(a, _req("filter_agg", close))
    for b in (far, other, nulls):
        assert not serving.same_result(a, _req("filter_agg", b))


@pytest.mark.parametrize("kind", ["filter_agg", "join_agg"])
def test_specs_compile_in_prepared_mode(kind):
    # one generated kernel serves every literal of a request kind
    from oxidelake_amd import autotune, plan
    from oxidelake_amd.backend import aggregate_kernel_source

    autotune.set_knob("jit.params", 1)
    try:
        if kind == "filter_agg":
            cols = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT32, True), (plan.FLOAT64, False)]
            srcs = {aggregate_kernel_source(cols, serving.filter_agg_spec(t), compile=t == 1)
                    for t in (1, 50, 500)}
        else:
            cols = [(plan.INT64, False), (plan.INT32, False), (plan.FLOAT64, False)]
            srcs = {aggregate_kernel_source(cols, serving.join_agg_spec(lo, lo + 9), probe_key=0,
                                            compile=lo == 0) for lo in (0, 100, 5000)}
        assert len(srcs) == 1 and "L[0]" in srcs.pop()
    finally:
        autotune.set_knob("jit.params", -1)


def test_vector_footprint_follows_the_pilot_path():
    # an index of at least 4 pilots scans its pilot into a score tile and
    # appends the rest to a list: nowhere near a score tile of every row
    big = serving.vector_scratch(20_000_000, 768, 16, 100)
    assert big < 200 << 20
    assert big > 16 * (1 << 20) * 4  # the pilot's scores (1M rows) are there
    # below 4 pilots the plain path scores every row (tiles of at most 2 GiB)
    small = serving.vector_scratch(3_000_000, 768, 16, 100)
    assert 16 * 3_000_000 * 4 <= small < 16 * 3_000_000 * 4 + (64 << 20)
    assert serving.vector_scratch(20_000_000, 768, 512, 1000) > serving.vector_scratch(20_000_000, 768, 256, 1000)

"""Test doubles around `oxidelake_amd.operator.GpuOperator`.

`OracleCpu` is the CPU path a caller hands the operator, over the oracle's
restatement of crates/oxidelake-device/src/cpu/kernels.rs. `FakeBackend`
claims to be a device and runs the oracle (the reference's `FakeGpu`,
crates/oxidelake-compute/src/operator.rs:1158-1327): it declines what it is
told to and counts what the operator asked of it, so the operator's host
logic (sub-batch cut, build once, pieces, fallback and continuation,
counters) is tested without a GPU.
"""
from __future__ import annotations

import pyarrow as pa

from oxidelake_amd import plan
from oxidelake_amd.backend import Unsupported
from oxidelake_amd.loader import device_type


class OracleCpu:
    def __init__(self, oracle):
        self.o = oracle

    def filter_project(self, batch, predicate, projection, names):
        return self.o.filter_project(batch, predicate, projection, names)

    def join_matches(self, left, left_key, build):
        table = build.host_state.get("oracle")
        if table is None:
            table = build.host_state["oracle"] = self.o.JoinTable(build.batch, build.key)
        return self.o.JoinMatches(left, left_key, table)

    def join_gather(self, left, build, matches, first, rows, projection):
        return self.o.hash_join_gather(left, build.batch, matches, first, rows, projection)

    def aggregator(self, schema_batch, spec):
        return self.o.GroupedAggregator(schema_batch, spec)

    def vector_distance(self, batch, column, query, metric, output_name):
        return self.o.vector_distance(batch, column, query, metric, output_name)


class _Stream:
    def synchronize(self):
        pass


class _Dev:
    def __init__(self, batch: pa.RecordBatch):
        self.batch = batch
        self.num_rows = batch.num_rows


def _mentions_float(e) -> bool:
    if isinstance(e, plan.Lit):
        return e.type in (plan.FLOAT32, plan.FLOAT64)
    if isinstance(e, plan.Bin):
        return _mentions_float(e.left) or _mentions_float(e.right)
    return False


class _AggState:
    def __init__(self, fake, like, spec):
        self.fake = fake
        self.g = fake.o.GroupedAggregator(like.batch, spec)

    def update(self, stream, dev):
        n = len(self.fake.aggregate_rows)
        if self.fake.declines_update is not None and n in self.fake.declines_update:
            self.fake.aggregate_rows.append(None)
            raise Unsupported("fake.aggregate", "this fake declines some batches")
        self.fake.aggregate_rows.append(dev.num_rows)
        self.g.update(dev.batch)

    def finish(self, stream):
        return _Dev(self.g.finish())


class FakeBackend:
    def __init__(self, oracle, float_predicates=True, declines_prepare=False, gathers_before_declining=None,
                 hash_join=True, aggregate=True, declines_update=None):
        self.o = oracle
        self.float_predicates = float_predicates
        self.declines_prepare = declines_prepare
        self.gathers_before_declining = gathers_before_declining
        self.hash_join_ok = hash_join
        self.aggregate_ok = aggregate
        self.declines_update = declines_update
        self.uploads = self.prepared = self.probes = self.joins = 0
        self.aggregate_rows: list = []

    def create_stream(self):
        return _Stream()

    def upload(self, stream, batch):
        for f in batch.schema:
            if not device_type(f.type):
                raise Unsupported("device.column_type", f"{f.type} columns are not device-resident")
        self.uploads += 1
        return _Dev(batch)

    def download(self, stream, dev):
        return dev.batch

    def supports_predicate(self, predicate):
        return self.float_predicates or not _mentions_float(predicate)

    def supports_hash_join(self):
        return self.hash_join_ok

    def supports_aggregate(self):
        return self.aggregate_ok

    def filter_project(self, stream, dev, predicate, projection, names=None):
        return _Dev(self.o.filter_project(dev.batch, predicate, projection, names))

    def prepare_join_build(self, stream, dev, key):
        if self.declines_prepare:
            raise Unsupported("backend.prepare_join_build", "this fake declines it")
        self.prepared += 1
        return self.o.JoinTable(dev.batch, key)

    def probe_join(self, stream, dev, key, table):
        self.probes += 1
        return self.o.JoinMatches(dev.batch, key, table)

    def hash_join(self, stream, left, right, table, matches, first, rows):
        done = self.joins
        self.joins += 1
        if self.gathers_before_declining is not None and done >= self.gathers_before_declining:
            raise Unsupported("fake.hash_join", "this fake declines join pieces after a few")
        return _Dev(self.o.hash_join_gather(left.batch, right.batch, matches, first, rows))

    def aggregate_state(self, stream, like, spec):
        return _AggState(self, like, spec)

    def vector_distance(self, stream, dev, column, query, metric, output_name):
        return _Dev(self.o.vector_distance(dev.batch, column, query, metric, output_name))

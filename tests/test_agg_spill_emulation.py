"""CPU emulation of the generated aggregation kernel's row routing with the
spill enabled (oxidelake_amd/csrc/aggregate_jit.h `aggregate_tiles`,
`spill_scatter` and `spill_fold`, and the spill scan in
oxidelake_amd/csrc/aggregate.hip), on the
input of tests/test_aggregate_gpu.py::test_spill_many_groups.

Each wave step is run with the kernel's SIMT semantics: the matching round
(lanes sharing the first selected lane's key fold into it), the block-table
probe in lock step (every lane reads its slot, then the compare-and-swaps
land in lane order), the per-wave `fill` flag read when the wave starts a
tile, spill buckets whose atomic position counter runs past their capacity
(rows past it take the HBM table), the histogram over partitions, the
exclusive scan that places each (partition, block) run, the scatter and the
fold. The waves of a block interleave in a random order at wave-step
granularity, a different order per seed. What is checked: every selected
row is counted exactly once (COUNT(*), and SUM/MAX of an integer column,
per group, against the oracle), and the scatter's positions fill the
record array exactly.
"""
import numpy as np
import pyarrow as pa
import pytest

from oxidelake_amd.plan import AggregateFunction as F
from oxidelake_amd.plan import Aggregate, AggregateSpec, Lit, col
from oxidelake_amd.plan import Comparison as Cmp
from oxidelake_amd import plan

LDS_LOG2 = 10
LDS_SLOTS = 1 << LDS_LOG2
LDS_PROBE = 8
SPILL_BUCKETS = 32
MAX_PARTS = 1024
EMPTY = None
I64MIN = -(2**63)


def lds_hash(key: int) -> int:
    u = key & (2**64 - 1)
    f = (u & 0xFFFFFFFF) ^ (u >> 32)
    return ((f * 0x9E3779B1) & 0xFFFFFFFF) >> (32 - LDS_LOG2)


def spill_test_input(n=1_000_000):
    # the data of test_spill_many_groups (seed 23)
    rng = np.random.default_rng(23)
    k = rng.integers(-150_000, 150_000, n).astype(np.int64)
    kmask = rng.random(n) < 0.01
    rng.uniform(0.5, 5, n)  # v: drawn to keep the stream aligned, not used here
    rng.random(n)
    w = rng.integers(0, 1000, n).astype(np.int32)
    return k, ~kmask, w


def emulate(k, kvalid, w, S, grid_cap, pct, seed):
    """Returns {key or None: [count, sum_w, max_w]} as the device would
    leave it, and the number of records the scatter placed."""
    n = len(k)
    tile = 256 * S
    ntiles = (n + tile - 1) // tile
    grid = min(ntiles, grid_cap)
    sel_all = w < 900
    # the key range of the valid keys (key_range_kernel) and the spill shape
    kmin, kmax = int(k[kvalid].min()), int(k[kvalid].max())
    count = kmax - kmin + 1
    # fold_log2: 4 aggregates (8-byte accumulators), a used byte, one
    # nullable input (SUM(v))
    slog2 = 14
    while slog2 > 5 and (1 << slog2) * (8 * 4 + 1 + 1) > 64 * 1024:
        slog2 -= 1
    nparts = (count + (1 << slog2) - 1) >> slog2
    assert nparts <= MAX_PARTS
    sper = (nparts + SPILL_BUCKETS - 1) // SPILL_BUCKETS
    tiles_per_block = (ntiles + grid - 1) // grid
    cap = max(64, (tiles_per_block * tile * pct // 100 + SPILL_BUCKETS - 1) // SPILL_BUCKETS)

    hbm = {}  # the HBM table: key -> [count, sum, max]

    def hbm_apply(key, c):
        acc = hbm.setdefault(key, [0, 0, -(2**63)])
        acc[0] += c[0]
        acc[1] += c[1]
        acc[2] = max(acc[2], c[2])

    rng = np.random.default_rng(seed)
    hist = np.zeros((nparts, grid), np.int64)
    buckets = []  # per block: 32 lists of records (g, count, sum, max)
    for b in range(grid):
        s_keys = [EMPTY] * (LDS_SLOTS + 2)
        s_acc = [[0, 0, -(2**63)] for _ in range(LDS_SLOTS + 2)]
        s_used = [0]
        s_bucket = [0] * SPILL_BUCKETS
        blk = [[] for _ in range(SPILL_BUCKETS)]
        for t in range(b, ntiles, grid):
            base = t * tile
            lim = min(n - base, tile)
            fill = {}
            # the 4 waves' steps in a random interleaving (each wave in order)
            order = np.repeat(np.arange(4), S)
            rng.shuffle(order)
            step_of = [0, 0, 0, 0]
            for wave in order:
                s = step_of[wave]
                step_of[wave] += 1
                if s == 0:
                    fill[wave] = s_used[0] < LDS_SLOTS // 2
                lanes = np.arange(64)
                j = s * 256 + wave * 64 + lanes
                inb = j < lim
                rows = base + np.minimum(j, lim - 1)
                sel = inb & sel_all[rows]
                kv = sel & kvalid[rows]
                key = np.where(kv, k[rows], 0)
                c = [[1, int(w[r]), int(w[r])] for r in rows]
                # matching round
                act = np.nonzero(sel)[0]
                merged = -1
                if act.size:
                    lead = act[0]
                    same = sel & (key == key[lead]) & (kv == kv[lead])
                    if same.sum() > 1:
                        m = np.nonzero(same)[0]
                        c[lead] = [int(same.sum()), sum(c[i][1] for i in m), max(c[i][2] for i in m)]
                        sel = sel & ~(same & (lanes != lead))
                        merged = lead
                # block-table probe in lock step
                lslot = np.full(64, -1)
                todo = []
                for i in np.nonzero(sel)[0]:
                    if not kv[i]:
                        lslot[i] = LDS_SLOTS
                        s_keys[LDS_SLOTS] = 0
                    else:
                        todo.append(i)
                h = {i: lds_hash(int(key[i])) for i in todo}
                for _ in range(LDS_PROBE):
                    if not todo:
                        break
                    cur = {i: s_keys[h[i]] for i in todo}  # every lane reads, then CAS in lane order
                    nxt = []
                    for i in todo:
                        kk = int(key[i])
                        if cur[i] == kk:
                            lslot[i] = h[i]
                            continue
                        if cur[i] is EMPTY:
                            if not fill[wave]:
                                continue  # not held
                            old = s_keys[h[i]]
                            if old is EMPTY:
                                s_keys[h[i]] = kk
                                s_used[0] += 1
                                lslot[i] = h[i]
                                continue
                            if old == kk:
                                lslot[i] = h[i]
                                continue
                        h[i] = (h[i] + 1) & (LDS_SLOTS - 1)
                        nxt.append(i)
                    todo = nxt
                # spill (single rows only: a record's COUNT(*) is implied 1 and
                # its Int32 values are 32 bits), else HBM, else the block table
                for i in np.nonzero(sel)[0]:
                    kk = int(key[i])
                    if lslot[i] < 0 and i != merged and kv[i] and kmin <= kk < kmin + count:
                        g = kk - kmin
                        part = g >> slog2
                        bucket = part // sper
                        pos = s_bucket[bucket]
                        s_bucket[bucket] += 1
                        if pos < cap:
                            assert c[i][0] == 1, "a spill record holds one row"
                            blk[bucket].append((g, *c[i]))
                            hist[part, b] += 1
                            continue
                    if lslot[i] >= 0:
                        acc = s_acc[lslot[i]]
                        acc[0] += c[i][0]
                        acc[1] += c[i][1]
                        acc[2] = max(acc[2], c[i][2])
                    else:
                        hbm_apply(kk if kv[i] else None, c[i])
        # the block table's flush
        for slot in range(LDS_SLOTS + 2):
            if s_keys[slot] is EMPTY:
                continue
            hbm_apply(None if slot == LDS_SLOTS else s_keys[slot], s_acc[slot])
        for c_ in range(SPILL_BUCKETS):
            assert len(blk[c_]) == min(s_bucket[c_], cap)
        buckets.append(blk)
    # exclusive scan over hist laid out [part][block]; scatter; fold
    flat = hist.reshape(-1)
    offs = np.concatenate([[0], np.cumsum(flat)])
    total = int(offs[-1])
    out = [None] * total
    for b in range(grid):
        for c_ in range(SPILL_BUCKETS):
            cur = {p: int(offs[p * grid + b]) for p in range(c_ * sper, min(nparts, (c_ + 1) * sper))}
            for rec in buckets[b][c_]:
                p = rec[0] >> slog2
                out[cur[p]] = rec
                cur[p] += 1
            for p in cur:
                assert cur[p] == offs[p * grid + b + 1], "a block's run of a partition was not filled exactly"
    assert all(r is not None for r in out)
    for p in range(nparts):
        lo, hi = int(offs[p * grid]), int(offs[(p + 1) * grid])
        for g, cnt, sm, mx in out[lo:hi]:
            assert g >> slog2 == p
            hbm_apply(kmin + g, [cnt, sm, mx])
    return hbm, total


def expected(k, kvalid, w):
    sel = w < 900
    want = {}
    for kk, v, ww in zip(k[sel], kvalid[sel], w[sel]):
        key = int(kk) if v else None
        acc = want.setdefault(key, [0, 0, -(2**63)])
        acc[0] += 1
        acc[1] += int(ww)
        acc[2] = max(acc[2], int(ww))
    return want


@pytest.fixture(scope="module")
def data():
    k, kv, w = spill_test_input()
    return k, kv, w, expected(k, kv, w)


@pytest.mark.parametrize("seed", [0, 1])
def test_failing_case_shape_conserves_every_row(data, seed):
    # the shape the generated kernel takes for that test: 8 rows per lane
    # (7 row registers + 2 per aggregate), one tile per block on 256 CUs x 4,
    # buckets of 64 records at agg.spill_pct = 1
    k, kv, w, want = data
    got, spilled = emulate(k, kv, w, S=8, grid_cap=1024, pct=1, seed=seed)
    assert spilled > 0
    assert got == want


@pytest.mark.parametrize("S,grid_cap,pct", [(4, 64, 1), (16, 200, 100), (8, 37, 5)])
def test_other_shapes_conserve_every_row(data, S, grid_cap, pct):
    # several tiles per block (the fill flag turns off part way), and spill
    # capacities from overflowing to ample
    k, kv, w, want = data
    n = 200_000
    k, kv, w = k[:n], kv[:n], w[:n]
    got, _ = emulate(k, kv, w, S=S, grid_cap=grid_cap, pct=pct, seed=S)
    assert got == expected(k, kv, w)

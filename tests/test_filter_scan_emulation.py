"""CPU emulation of the generated filter kernel's in-tile compaction
(oxidelake_amd/csrc/filter_jit.h, filter_tile steps 2-4), for every rows per
lane S the skeleton takes (1-32).

A tile is 256 threads x S rows; row s*256 + tid sits on thread tid, i.e. on
lane tid % 64 of wave tid // 64. Step 2 stores each (step, wave)'s ballot and
its popcount at index s*4 + wave; step 3 scans the S*4 counts on one wave,
PL = ceil(S*4/64) consecutive counts per lane; step 4 gives a kept row the
output position s_off[s*4 + wave] + popcount(ballot & lanes below it).

The property checked is what the compaction must give, independent of how
the kernel computes it: the kept rows of a tile land on positions 0..kept-1,
each exactly once, in row order (the order `filter_record_batch` keeps,
crates/oxidelake-device/src/cpu/kernels.rs:144-152). S = 20 once faulted on
the device because counts past 64 were never scanned; this emulation runs
the same index arithmetic on the CPU for every S.
"""
import numpy as np
import pytest

BLOCK, WAVES, LANES = 256, 4, 64


def emulate_tile(kept: np.ndarray, S: int) -> np.ndarray:
    """Output position of every row of one tile (-1 where not kept), computed
    as filter_tile does. `kept` has 256 * S flags in row order."""
    NC = S * WAVES
    PL = (NC + LANES - 1) // LANES
    flags = kept.reshape(S, WAVES, LANES)  # [step][wave][lane]
    # step 2: ballots (bit = lane) and their popcounts, at index s*4 + wave
    weights = np.uint64(1) << np.arange(LANES, dtype=np.uint64)
    masks = (flags.astype(np.uint64) * weights).sum(axis=2, dtype=np.uint64).reshape(NC)
    cnt = flags.sum(axis=2).reshape(NC).astype(np.uint32)
    # step 3: one wave, PL consecutive counts per lane, inclusive wave scan
    v = np.zeros((LANES, PL), dtype=np.uint32)
    for lane in range(LANES):
        for k in range(PL):
            c = lane * PL + k
            if c < NC:
                v[lane, k] = cnt[c]
    lane_sum = v.sum(axis=1, dtype=np.uint32)
    incl = np.cumsum(lane_sum, dtype=np.uint32)  # wave_incl_scan
    off = np.full(NC, 0xFFFFFFFF, dtype=np.uint64)
    for lane in range(LANES):
        run = int(incl[lane] - lane_sum[lane])
        for k in range(PL):
            c = lane * PL + k
            if c < NC:
                off[c] = run
            run += int(v[lane, k])
    assert (off != 0xFFFFFFFF).all(), "a (step, wave) offset was never written"
    # step 4: q = s_off[s*4 + wave] + popcount(mask & below)
    q = np.full(BLOCK * S, -1, dtype=np.int64)
    for s in range(S):
        for w in range(WAVES):
            m = int(masks[s * WAVES + w])
            for lane in range(LANES):
                if (m >> lane) & 1:
                    below = m & ((1 << lane) - 1)
                    q[s * BLOCK + w * LANES + lane] = int(off[s * WAVES + w]) + bin(below).count("1")
    return q


@pytest.mark.parametrize("S", list(range(1, 33)))
def test_tile_positions_are_the_kept_rows_in_order(S):
    rng = np.random.default_rng(1000 + S)
    n = BLOCK * S
    cases = [np.zeros(n, bool), np.ones(n, bool)]
    for p in (0.01, 0.155, 0.5, 0.97):
        cases.append(rng.random(n) < p)
    lim = int(rng.integers(1, n))  # a short last tile: rows at or past lim are not kept
    part = rng.random(n) < 0.4
    part[lim:] = False
    cases.append(part)
    for kept in cases:
        q = emulate_tile(kept, S)
        want = np.where(kept, np.cumsum(kept) - 1, -1)
        np.testing.assert_array_equal(q, want)


def test_rows_per_lane_fit_the_kept_bits():
    # the kept flags of a lane are the bits of one 32-bit register, and the
    # library refuses a rows-per-lane knob above that (tests/test_abi.py)
    import re
    from pathlib import Path

    src = (Path(__file__).resolve().parents[1] / "oxidelake_amd" / "csrc" / "filter_jit.h").read_text()
    assert int(re.search(r"kMaxRowsPerLane = (\d+);", src).group(1)) == 32
    assert "uint32_t selbits" in src

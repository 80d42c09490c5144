"""Project autotuner: picks kernel launch shapes by timing them on the device.

Knobs live in the library (oxr_tune_set). The tuner sweeps the instantiated
candidates of a knob group on a representative batch, timing each with the
library's HIP-event kernel profiler (oxr_profile_*), keeps the fastest and
records the choice in a JSON cache keyed by device and workload, so a later
process applies it without re-timing.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import Callable, Iterable, Sequence

from ._capi import lib

CACHE = Path(os.environ.get("OXR_TUNE_CACHE", Path.home() / ".cache" / "oxidelake_amd" / "tune.json"))

FILTER_SHAPES = [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2), (4, 1)]  # interpreter kernel
FILTER_ROWS = [8, 16, 24, 32]  # generated kernels: rows per lane
# the staged filter kernel: (waves per workgroup, chunks of 4 rows per lane,
# kept rows its LDS stage holds per tile)
FILTER_STAGED = [(8, 8, 3072), (8, 6, 2560), (4, 8, 1536)]
AGG_GRIDS = [2, 4, 8, 16]
# vector search, large batches: (vector.gemm, vector.pilot_rows) -- the GEMM
# tile or the resident/streamed kernels, and how many first rows bound the
# rest (a small pilot costs little selection but lets more keys through the
# bound; a large one the reverse)
VECTOR_LARGE = [(1, 1 << 19), (1, 3 << 18), (1, 1 << 20), (1, 3 << 19), (1, 1 << 21), (1, 1 << 22), (0, 1 << 21)]
# join + aggregate over a wide skewed key range: (agg.spill_pct, agg.hot_grid)
# -- spill bucket size in percent of a block's rows, hot-kernel workgroups
# per CU
JOIN_SPILL = [(40, 4), (40, 8), (60, 4), (60, 8), (90, 4), (90, 8)]
# batched scalar aggregations: (agg.multi_rows, agg.multi_grid)
AGG_BATCH = [(2, 2), (4, 1), (4, 2), (4, 4), (8, 2)]


def set_knob(name: str, value: int) -> None:
    """Sets a library knob (a negative value resets it). Raises PlanError for
    an unknown knob or a value its kernels do not take."""
    from .backend import check

    check(lib.oxr_tune_set(name.encode(), int(value)))


def get_knob(name: str) -> int:
    v = C.c_int64()
    lib.oxr_tune_get(name.encode(), C.byref(v))
    return int(v.value)


def kernel_ms(kernel: str, run: Callable[[], object], reps: int) -> float:
    """Average device time of `kernel` over `reps` calls of `run`."""
    run()  # warm (allocations, code objects)
    lib.oxr_profile_reset()
    lib.oxr_profile_enable(1)
    try:
        for _ in range(reps):
            run()
    finally:
        lib.oxr_profile_enable(0)
    ms, n = C.c_double(), C.c_uint64()
    lib.oxr_profile_get(kernel.encode(), C.byref(ms), C.byref(n))
    lib.oxr_profile_reset()
    return ms.value / max(1, n.value)


def device_ms(kernels: Sequence[str], run: Callable[[], object], reps: int) -> float:
    """Device time of one call of `run`: every launch of the named kernels
    over `reps` calls, summed, divided by `reps` (a call may launch each
    several times)."""
    run()  # warm (allocations, code objects)
    lib.oxr_profile_reset()
    lib.oxr_profile_enable(1)
    try:
        for _ in range(reps):
            run()
    finally:
        lib.oxr_profile_enable(0)
    total = 0.0
    for k in kernels:
        ms, n = C.c_double(), C.c_uint64()
        lib.oxr_profile_get(k.encode(), C.byref(ms), C.byref(n))
        total += ms.value
    lib.oxr_profile_reset()
    return total / max(1, reps)


def launches(kernel: str) -> int:
    """Launches of `kernel` recorded by the profiler since its last reset."""
    ms, n = C.c_double(), C.c_uint64()
    lib.oxr_profile_get(kernel.encode(), C.byref(ms), C.byref(n))
    return int(n.value)


def sweep(kernel, knobs: Sequence[str], candidates: Iterable[Sequence[int]],
          run: Callable[[], object], reps: int = 3) -> dict:
    """Times every candidate setting of `knobs`, applies the fastest.
    `kernel`: one kernel name (its average launch), or a sequence of names
    (their device time per call of `run`)."""
    timings = {}
    for cand in candidates:
        for k, v in zip(knobs, cand):
            set_knob(k, v)
        timings[",".join(map(str, cand))] = (kernel_ms(kernel, run, reps) if isinstance(kernel, str)
                                             else device_ms(kernel, run, reps))
    best = min(timings, key=timings.get)
    for k, v in zip(knobs, best.split(",")):
        set_knob(k, int(v))
    return {"knobs": dict(zip(knobs, map(int, best.split(",")))), "ms": timings}


def tune_filter(run: Callable[[], object], reps: int = 3) -> dict:
    """The generated filter's two kernels on the device and table at hand:
    the tile kernel at each rows per lane (the staged kernel off), and the
    staged kernel (kept rows parked in LDS) at each shape of FILTER_STAGED,
    which the query takes while its kept fraction fits the stage."""
    knobs = ["filter.staged", "filter.rows", "filter.staged_waves", "filter.chunks", "filter.stage_rows"]
    cands = [(0, r, -1, -1, -1) for r in FILTER_ROWS] + [(1, 24, w, c, st) for w, c, st in FILTER_STAGED]
    return sweep("filter_project", knobs, cands, run, reps)


def tune_filter_interpreter(run: Callable[[], object], reps: int = 3) -> dict:
    """Launch shape of the interpreter kernel (filter.jit = 0)."""
    return sweep("filter_project_vm", ["filter.v", "filter.g"], FILTER_SHAPES, run, reps)


def tune_aggregate(run: Callable[[], object], reps: int = 3) -> dict:
    return sweep("aggregate", ["agg.grid"], [(g,) for g in AGG_GRIDS], run, reps)


def tune_vector_large(run: Callable[[], object], reps: int = 3) -> dict:
    """Scan kernel and pilot size of a large-batch vector search: the whole
    search's device time (scans, selections, merges, rerank)."""
    return sweep(("vector_scan", "vector_select", "vector_merge", "vector_rerank"),
                 ["vector.gemm", "vector.pilot_rows"], VECTOR_LARGE, run, reps)


def tune_join_aggregate(run: Callable[[], object], reps: int = 2) -> dict:
    """Spill bucket size and hot-kernel grid of a fused join + aggregate
    update over a wide skewed key range: every pass of the update."""
    return sweep(("aggregate", "aggregate_hot", "aggregate_hot_map", "aggregate_hot_reduce", "aggregate_spill",
                  "key_range"), ["agg.spill_pct", "agg.hot_grid"], JOIN_SPILL, run, reps)


def tune_agg_batch(run: Callable[[], object], reps: int = 3) -> dict:
    """Rows per lane and workgroups per CU of the batched-request kernels.
    `run` must prepare its batch object itself (the grid is fixed when a
    batch is prepared)."""
    return sweep(("aggregate_batch",), ["agg.multi_rows", "agg.multi_grid"], AGG_BATCH, run, reps)


def reset(knobs: Iterable[str]) -> None:
    for k in knobs:
        set_knob(k, -1)


def save(key: str, result: dict, path: Path = CACHE) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    data = json.loads(path.read_text()) if path.exists() else {}
    data[key] = result
    path.write_text(json.dumps(data, indent=1))


def apply(key: str, path: Path = CACHE) -> bool:
    """Applies a cached choice; False when none is cached for `key`."""
    if not path.exists():
        return False
    entry = json.loads(path.read_text()).get(key)
    if not entry:
        return False
    for k, v in entry["knobs"].items():
        set_knob(k, v)
    return True

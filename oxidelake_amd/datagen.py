"""Seeded synthetic tables for the benchmark workloads.

Device-side generation (torch on the GPU) for the HBM-resident full-size
tables, and conversion of the same tensors to pyarrow RecordBatches for the
CPU oracle at test sizes. The reference's own demo generator is SplitMix64
(crates/oxidelake-storage/src/datagen.rs:89-95); `splitmix64` below restates
it for the pinned demo-table fixture.
"""
from __future__ import annotations

import numpy as np
import pyarrow as pa

from . import plan

MASK64 = (1 << 64) - 1


def splitmix64(state: int) -> tuple[int, int]:
    """datagen.rs:89-95: returns (new_state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return state, z ^ (z >> 31)


def pack_validity_torch(valid):
    """Bool tensor (n) -> LSB-first bitmap bytes, padded to whole 32-bit words."""
    import torch

    n = valid.shape[0]
    words = (n + 31) // 32
    pad = words * 32 - n
    v = valid.to(torch.uint8)
    if pad:
        v = torch.cat([v, torch.zeros(pad, dtype=torch.uint8, device=v.device)])
    v = v.view(-1, 8).to(torch.int32)
    w = torch.tensor([1, 2, 4, 8, 16, 32, 64, 128], dtype=torch.int32, device=v.device)
    return (v * w).sum(dim=1).to(torch.uint8).contiguous()


def config2_torch(n: int, seed: int = 42, device: str = "cuda", null_frac: float = 0.05):
    """Config 2's table: a Int64 U[0,2^40), b Int32 U[0,1000), c Float32 N(0,1)
    with `null_frac` NULLs, d Float64 lognormal(0,1). Returns a dict of
    tensors: a, b, c, c_valid (bool), c_bits (bitmap), d."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    a = torch.randint(0, 1 << 40, (n,), dtype=torch.int64, device=device, generator=g)
    b = torch.randint(0, 1000, (n,), dtype=torch.int32, device=device, generator=g)
    c = torch.randn(n, dtype=torch.float32, device=device, generator=g)
    c_valid = torch.rand(n, device=device, generator=g) >= null_frac
    d = torch.randn(n, dtype=torch.float64, device=device, generator=g).exp_()
    return {"a": a, "b": b, "c": c, "c_valid": c_valid, "c_bits": pack_validity_torch(c_valid), "d": d}


def config2_query(b_threshold: int = 50, c_threshold: float = 0.5, modulus: int = 7):
    """SELECT a, b*2+1, c*d FROM t WHERE b < T AND c > 0.5 OR a % 7 = 3
    over columns (a, b, c, d) = (0, 1, 2, 3). Returns (predicate, projection)."""
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


def config2_arrow(t: dict) -> pa.RecordBatch:
    """The same table on the host, as a RecordBatch (a, b, c, d)."""
    a = t["a"].cpu().numpy()
    b = t["b"].cpu().numpy()
    c = t["c"].cpu().numpy()
    cv = t["c_valid"].cpu().numpy()
    d = t["d"].cpu().numpy()
    return pa.record_batch({
        "a": pa.array(a, pa.int64()),
        "b": pa.array(b, pa.int32()),
        "c": pa.array(c, pa.float32(), mask=~cv),
        "d": pa.array(d, pa.float64()),
    })


def config2_wrap(backend, t: dict):
    """Config 2's tensors as a device batch (no copy)."""
    return backend.wrap([
        (t["a"], None, plan.INT64),
        (t["b"], None, plan.INT32),
        (t["c"], t["c_bits"], plan.FLOAT32),
        (t["d"], None, plan.FLOAT64),
    ], names=["a", "b", "c", "d"])


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

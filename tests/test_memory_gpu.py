"""GPU: the trait's raw memory methods and capability hooks through the C ABI
(crates/oxidelake-device/src/backend.rs:52-79, 182-197)."""
import numpy as np
import pytest

from oxidelake_amd.backend import ExecutionError

pytestmark = pytest.mark.gpu


def test_device_and_pinned_buffers_round_trip(backend, stream):
    n = (1 << 20) + 13  # not a multiple of any transfer granule
    dev = backend.alloc_device(n)
    host_in = backend.alloc_pinned_host(n)
    host_out = backend.alloc_pinned_host(n)
    assert len(dev) == n and host_in.class_ == "pinned"
    # alloc_device is zeroed
    backend.copy_d2h(stream, dev, host_out)
    stream.synchronize()
    assert not host_out.array().any()
    data = np.random.default_rng(5).integers(0, 256, n, dtype=np.uint8)
    host_in.array()[:] = data
    backend.copy_h2d(stream, host_in, dev)
    backend.copy_d2h(stream, dev, host_out)
    stream.synchronize()
    assert np.array_equal(host_out.array(), data)
    # a copy past the destination is refused before anything moves
    big = backend.alloc_pinned_host(n + 1)
    with pytest.raises(ExecutionError):
        backend.copy_h2d(stream, big, dev)
    with pytest.raises(ExecutionError):
        backend.copy_d2h(stream, dev, big)
    for b in (dev, host_in, host_out, big):
        b.free()


def test_capability_hooks(backend):
    assert backend.supports_hash_join() and backend.supports_aggregate()
    info = backend.memory_info()
    assert info.total_bytes > 200 * 2**30 and 0 < info.free_bytes <= info.total_bytes and info.pinned_supported


@pytest.mark.parametrize("staging,chunk_kib", [(0, 4096), (1, 4096), (1, 64)])
def test_upload_download_through_the_staging_ring(backend, stream, staging, chunk_kib):
    # every column type, NULLs, a slice at an odd offset, buffers of many
    # chunks with ragged ends: the same batch comes back either way
    import pyarrow as pa

    from oxidelake_amd import autotune

    rng = np.random.default_rng(17)
    n = 300_007
    emb = pa.FixedSizeListArray.from_arrays(pa.array(rng.normal(size=n * 8).astype(np.float32)), 8)
    batch = pa.record_batch({
        "i64": pa.array(rng.integers(-2**62, 2**62, n)),
        "i32": pa.array(rng.integers(-2**31, 2**31, n).astype(np.int32), mask=rng.random(n) < 0.1),
        "f32": pa.array(rng.normal(size=n).astype(np.float32)),
        "f64": pa.array(rng.normal(size=n), mask=rng.random(n) < 0.3),
        "emb": emb,
    })
    autotune.set_knob("transfer.staging", staging)
    autotune.set_knob("transfer.chunk_kib", chunk_kib)
    try:
        for b in (batch, batch.slice(3, n - 10), batch.slice(0, 0), batch.slice(5, 1)):
            got = backend.download(stream, backend.upload(stream, b))
            assert got.num_rows == b.num_rows
            for name in b.schema.names:
                assert got.column(name).equals(b.column(name)), name
    finally:
        autotune.set_knob("transfer.staging", -1)
        autotune.set_knob("transfer.chunk_kib", -1)

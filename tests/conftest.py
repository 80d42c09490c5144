import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def backend():
    """The ROCm backend on device 0. Fails (never skips) when the library or
    the GPU is missing: a GPU test must not pass without running the HIP path."""
    from oxidelake_amd.backend import RocmBackend

    return RocmBackend(0)


@pytest.fixture()
def stream(backend):
    s = backend.create_stream()
    yield s
    s.synchronize()

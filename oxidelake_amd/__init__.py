"""MI355X (gfx950) backend for OxideLake's four device operators.

The product is liboxide_rocm.so (C ABI: include/oxide_rocm.h). This package
binds it (`backend.RocmBackend`, mirroring the reference's `GpuBackend`
trait), holds the plan vocabulary (`plan`), the per-batch operator contract
(`operator.GpuOperator`) and the build recipe (`build`).
"""
from . import plan  # noqa: F401

__all__ = ["plan"]

"""pimdal_amd — PIMDAL's DB-operator hot path on one MI355X (gfx950).

The product is libpimdal_gpu.so: hand-written HIP kernels behind the C ABI
declared in include/pimdal_gpu.h.  This package binds that ABI with ctypes
for tests, benchmarks and Python callers.
"""
from ._lib import PimdalError, load  # noqa: F401
from .device import Context, DeviceArray  # noqa: F401
from . import operators, datagen, serving, ingest  # noqa: F401

__all__ = ["Context", "DeviceArray", "PimdalError", "load", "operators", "datagen", "serving", "ingest"]

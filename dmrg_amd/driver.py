"""ctypes declarations of the sweep-driver entry points (include/dmrgx_driver.h)."""
from __future__ import annotations

DRIVER_EXPORTS: list[str] = []


def declare(lib):
    return lib

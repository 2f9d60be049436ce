"""MI355X-native explicit step of PeriDEM (bond-based peridynamics + DEM contact).

The product is the HIP library libperidem_gpu.so behind the C-ABI of
include/peridem_gpu.h.  This package holds its sources (csrc/), the in-tree
build (build.py), the ctypes view of the ABI (abi.py) and the synthetic
setup generator (gen.py).
"""
from .abi import Engine, PDError, gpu_engine, load_gpu_lib  # noqa: F401

__all__ = ["Engine", "PDError", "gpu_engine", "load_gpu_lib"]

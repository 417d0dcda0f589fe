"""muesli4_amd -- MI355X (gfx950) skeleton library: Python binding.

The product is the C ABI in ``include/msl_hip.h`` implemented by
``muesli4_amd/lib/libmsl_hip.so`` (hand-written HIP kernels) and the
msl::DA / DM / DC<T> C++ headers in ``muesli4_amd/include``.  This module
binds the C ABI with ctypes for the tests and the benchmark.  It never falls
back to a CPU implementation: when the library is missing, every call raises.
"""
from __future__ import annotations

from .hip import (  # noqa: F401
    MslError,
    DeviceArray,
    lib,
    lib_path,
    init,
    synchronize,
    device_info,
)

__all__ = ["MslError", "DeviceArray", "lib", "lib_path", "init", "synchronize", "device_info"]

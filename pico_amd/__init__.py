"""pico_amd -- MI355X-native (gfx950) k-core decomposition engine.

The product is libkcore.so (HIP kernels + C ABI, include/kcore.h) plus host
C++ drop-ins for the reference executables; this package is the Python view
of that ABI.  See pico_amd.kcore.
"""
from .kcore import (  # noqa: F401
    METHODS, Context, DeviceBatch, DeviceGraph, HostBatch, HostGraph, KCoreError, Server,
    core_checksum, lib, ref_peel, vgraph_decompose,
)

__version__ = "0.1.0"

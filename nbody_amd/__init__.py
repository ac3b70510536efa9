"""MI355X-native N-body force engine: `hip` (direct) and `hip_bh` (Barnes-Hut)
engines behind the drons/nbody engine interface.

Native code lives in libnbody_hip.so (gfx950 kernels, include/nbody_hip.h) and
libnbody.so (C++ engines, solvers, bodies, include/nbody_host.h); this package
only loads them.
"""
from .api import (fake_memory, log_count, log_last, log_quiet, memory, nbody_create_engine,
                  nbody_create_solver, nbody_data, nbody_data_stream, nbody_data_stream_reader, nbody_engine,
                  nbody_solver)
from .native import device_count, hip_lib, host_lib

__all__ = [
    "nbody_data", "nbody_engine", "nbody_solver", "memory", "fake_memory",
    "nbody_data_stream", "nbody_data_stream_reader",
    "nbody_create_engine", "nbody_create_solver", "log_count", "log_last", "log_quiet",
    "device_count", "hip_lib", "host_lib",
]

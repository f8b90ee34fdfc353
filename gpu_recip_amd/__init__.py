"""gpu_recip_amd -- MI355X-native PME reciprocal-space engine.

The product is libpme_hip.so (HIP kernels for gfx950 behind the C ABI in
include/pme_recip.h).  This package holds its sources (csrc/), the build
script, a ctypes binding and a Python mirror of the reference's
CudaPMERecip / CudaEnergyVirial interface, plus the synthetic system
generator used by the tests and the benchmark.
"""
from ._lib import PmeError  # noqa: F401

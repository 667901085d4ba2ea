"""MI355X (gfx950) time-stepping core of the MPulse nonlinear Maxwell solver.

The compute path is libmpulse_hip.so (hand-written HIP kernels behind a C ABI, see
include/mpulse_hip.h); this package only holds its ctypes binding and workload builders.
"""
from . import abi, configs  # noqa: F401

"""MI355X implementation of WAM Cycle 7's per-step hot path (IMPLSCH + PROPAGS).

The compute path is libwamgpu.so (hand-written HIP for gfx950 behind the
C-ABI in include/wam_gpu.h); this package holds its ctypes binding, the host
mirror of the reference's call sequence, and seeded synthetic workloads.
"""
from .model import (WamError, WamGpu, WamMultiGrid, boundary_append, boundary_read, boundary_read_header,  # noqa: F401
                    boundary_write_header, map_write, restart_read, restart_write)
from . import workloads  # noqa: F401

"""MI355X (gfx950) implementation of SparkleDock's per-step GSO hot path.

The compute path is libsdgso.so (HIP kernels behind a C-ABI, include/sdgso.h);
`sdgso` is its ctypes binding, `workloads` builds the seeded synthetic inputs.
"""
from . import sdgso  # noqa: F401

__all__ = ["sdgso"]

"""MI355X LBFS + IBM + spring-network membrane time step (HIP, gfx950).

Python side: a thin ctypes binding over the C ABI of liblbfs_amd.so
(include/lbfs_abi.h), the seeded benchmark workloads, and the build script.
"""
from . import abi  # noqa: F401

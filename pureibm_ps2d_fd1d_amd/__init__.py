"""MI355X-native per-time-step fluid solve of PUReIBM-PS2D-FD1D.

The product is the HIP library ``_build/libpureibm_gpu.so`` behind the C-ABI
of ``include/pureibm_gpu.h`` and the Fortran host library
``_build/libpib_host.so``; this package holds their sources, their build and
a thin Python host (``solver``) over the same C-ABI.
"""
from . import _abi  # noqa: F401

__all__ = ["_abi"]

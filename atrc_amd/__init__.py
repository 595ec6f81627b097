"""atrc_amd — MI355X (gfx950) path tracer that stands in for Atrc's CPU `pt`.

The product is the C ABI of ``include/atrc_gpu.h`` implemented by
``libatrc_gpu.so`` (HIP kernels + host-side scene build) and the C++ host in
``atrc_amd/host`` (Atrc JSON config, factory, ``gpu_pt`` renderer, CLI).
This Python package holds the ctypes mirror of the ABI (``capi``), the seeded
scene generators (``scenes``) and the build script (``build``).
"""
from . import capi, scenes  # noqa: F401

__all__ = ["capi", "scenes"]

"""anacapa_amd — MI355X (gfx950) wavefront Path+NEE backend for the Anacapa renderer.

The product is the C-ABI library ``anacapa_amd/lib/libanacapa_hip.so`` (C++ host code plus
hand-written HIP kernels, declared in ``include/anacapa_hip.h``).  This Python package only
loads it (``_abi``), builds the procedural workloads (``scenes``) and mirrors the
reference's IIntegrator calls (``integrator``) for tests and the benchmark.
"""
from . import _abi, scenes  # noqa: F401

__all__ = ["_abi", "scenes"]

"""MI355X-native per-energy NEGF solve (gfx950 HIP behind a C ABI).

``capi``        -- ctypes binding of the C ABI (include/negf_capi.h)
``hamiltonian`` -- host-side tight-binding wire builder (inputs of the solve)
``build``       -- hipcc / amdflang build recipe for the in-tree libraries
"""
from . import hamiltonian  # noqa: F401

__all__ = ["capi", "hamiltonian", "build"]

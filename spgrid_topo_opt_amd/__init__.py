"""MI355X-native narrow-band SPGrid elasticity solve and compliance sensitivity.

Drop-in for the reference's `fem_solve(void*, void*)` (topo_opt/fem_interface.h)
plus a pointer-level C ABI (include/spgrid_topo_opt_amd.h) used by the Python
driver.  Hand-written HIP kernels for gfx950; no CPU fallback.
"""
from . import workloads  # noqa: F401

__all__ = ["workloads"]

"""MI355X-native (gfx950) caustic-design hot path of dylanmsu/poisson_caustic_design.

The product is the C-ABI shared library ``libcaustic_amd.so`` (include/cd_api.h)
with the C++ drop-ins ``solver.h`` / ``caustic_design.h``.  This package holds
its sources (csrc/), the build script and a ctypes mirror (capi) of the
reference's interfaces used by tests and benchmarks.
"""
from . import capi  # noqa: F401
from .capi import CausticDesign, Context, poisson_solver  # noqa: F401

__all__ = ["capi", "CausticDesign", "Context", "poisson_solver"]

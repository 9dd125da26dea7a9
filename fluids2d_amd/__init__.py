"""fluids2d_amd -- MI355X-native (gfx950) time-stepping hot path of Fluids2d.

Drop-in for the reference's ``Param`` / ``Model`` / integrator / ``State``
API on the GPU: WENO flux reconstruction, RK3 tendency/update kernels and the
masked Poisson/Helmholtz solves run as hand-written HIP kernels
(libfluids2d_amd.so, C ABI in include/fluids2d_amd.h).
"""
from .param import Param
from .model import Model
from . import tools

__all__ = ["Param", "Model", "tools"]
__version__ = "0.1.0"

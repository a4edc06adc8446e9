"""ikarus_amd — MI355X-native hot path of Ikarus's nonlinear mechanics (hex8 NonLinearElastic / EAS
element residual and tangent, block-CSR scatter, PCG inside NewtonRaphson under LoadControl).

The compute lives in libikarus_amd.so (hand-written HIP for gfx950) behind the C ABI of
include/ikarus_amd.h; this package is the host-side mirror of the reference operator API.
"""
from . import _capi as capi  # noqa: F401
from .assembler import (FERequirement, LoadControl, NewtonRaphson, PathFollowing,  # noqa: F401
                        SparseFlatAssembler, Device, nr_settings, solver_opts)
from . import problems  # noqa: F401

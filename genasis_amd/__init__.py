"""MI355X-native finite-volume step of the GenASiS Fluid_P_I fluid.

The arithmetic lives in the HIP library lib/libgx_fluid.so (C-ABI in
include/gx_fluid.h); this package holds the host-side mirror of the
reference's chart, integrator and step interfaces.
"""
from .chart import Chart  # noqa: F401
from .fluid import FluidParameters, Step_RK2_C_ASC  # noqa: F401
from .integrator import Integrator_C_PS, IntegratorParameters  # noqa: F401

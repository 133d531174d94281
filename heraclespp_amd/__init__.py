"""heraclespp_amd: the HERACLES++ hydro step on MI355X (gfx950).

HIP kernels (heraclespp_amd/csrc) behind a C-ABI (include/hclpp_hip.h,
libhclpp_hip.so), with a host-side mirror of the reference operator
interfaces (hydro.py) and the time-loop driver (solver.py).
"""

from .hydro import (DistributedBoundaryCondition, DomainError, Grid, HancockExtrapolationReconstruction,
                    IExtrapolationReconstruction, IFaceReconstruction, IGodunovScheme, IHydroReconstruction,
                    InternalMassGravity, LimitedLinearReconstruction, MUSCLHancockHydroReconstruction, PerfectGas,
                    PointMassGravity, Range, RiemannBasedGodunovScheme, UniformGravity, conv_cons_to_prim,
                    conv_prim_to_cons, factory_face_reconstruction, factory_godunov_scheme, integrate,
                    make_internal_mass_gravity, make_point_mass_gravity, make_uniform_gravity, minimum_internal_energy,
                    pressure_fix, regular_nodes, time_step)
from .solver import HydroSolver

__all__ = [
    "DistributedBoundaryCondition", "DomainError", "Grid", "HancockExtrapolationReconstruction",
    "IExtrapolationReconstruction", "IFaceReconstruction", "IGodunovScheme", "IHydroReconstruction",
    "InternalMassGravity", "make_internal_mass_gravity", "pressure_fix", "LimitedLinearReconstruction", "MUSCLHancockHydroReconstruction", "PerfectGas", "PointMassGravity", "Range",
    "RiemannBasedGodunovScheme", "UniformGravity", "conv_cons_to_prim", "conv_prim_to_cons",
    "factory_face_reconstruction", "factory_godunov_scheme", "integrate", "make_point_mass_gravity",
    "make_uniform_gravity", "minimum_internal_energy", "regular_nodes", "time_step", "HydroSolver",
]

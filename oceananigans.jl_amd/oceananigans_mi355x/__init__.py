"""Host-side mirror of the Oceananigans interface for the MI355X NonhydrostaticModel path.

The names, arguments and defaults follow the reference
(``NonhydrostaticModel(grid; advection, closure, buoyancy, coriolis, tracers, boundary_conditions)``,
``set!``, ``time_step!``, ``Simulation``, ``run!``); Julia's ``!`` functions are spelled with a
trailing underscore (``set_``, ``time_step_``, ``run_``).  All arithmetic runs in the HIP library
(``liboceananigans_mi355x.so``) through its C ABI; this module only builds the configuration,
moves host arrays and sequences calls.
"""
from ._abi import OmError, load_library  # noqa: F401
from .grids import Bounded, Flat, GridFittedBottom, ImmersedBoundaryGrid, MI355X, Periodic, RectilinearGrid  # noqa: F401
from .physics import (  # noqa: F401
    AnisotropicMinimumDissipation,
    BuoyancyTracer,
    Centered,
    ColumnwiseTridiagonalPreconditioner,
    ConjugateGradientPoissonSolver,
    DiagonallyDominantPreconditioner,
    FFTBasedPoissonSolver,
    FourierTridiagonalPoissonSolver,
    FieldBoundaryConditions,
    FluxBoundaryCondition,
    FPlane,
    GradientBoundaryCondition,
    LinearEquationOfState,
    ScalarDiffusivity,
    SeawaterBuoyancy,
    SmagorinskyLilly,
    ValueBoundaryCondition,
    WENO,
)
from .models import NonhydrostaticModel, build_config, set_, time_step_  # noqa: F401
from .simulations import Simulation, run_  # noqa: F401

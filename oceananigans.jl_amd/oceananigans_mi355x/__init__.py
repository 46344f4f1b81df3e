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
    LearnedPreconditioner,
    LinearEquationOfState,
    ScalarDiffusivity,
    SeawaterBuoyancy,
    SmagorinskyLilly,
    ValueBoundaryCondition,
    WENO,
)
from .models import Ensemble, NonhydrostaticModel, build_config, set_, time_step_  # noqa: F401
from .simulations import (  # noqa: F401
    Callback,
    Checkpointer,
    IterationInterval,
    NaNChecker,
    Simulation,
    TimeInterval,
    TimeStepWizard,
    cell_advection_timescale,
    cell_diffusion_timescale,
    conjure_time_step_wizard_,
    load_checkpoint_state,
    prognostic_state,
    restore_prognostic_state_,
    run_,
    set_simulation_,
)

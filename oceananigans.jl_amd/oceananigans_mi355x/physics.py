"""Physics options with the reference's names and defaults.

WENO / Centered               src/Advection/weno_reconstruction.jl:107-140, centered_reconstruction.jl:11-28
AnisotropicMinimumDissipation src/TurbulenceClosures/turbulence_closure_implementations/anisotropic_minimum_dissipation.jl:136-146
SmagorinskyLilly              .../Smagorinskys/lilly_coefficient.jl:109-114
ScalarDiffusivity             src/TurbulenceClosures/turbulence_closure_implementations/scalar_diffusivity.jl
BuoyancyTracer / SeawaterBuoyancy / LinearEquationOfState  src/BuoyancyFormulations/*.jl
FPlane                        src/Coriolis/f_plane.jl:30-51
boundary conditions           src/BoundaryConditions/boundary_condition.jl, field_boundary_conditions.jl
"""
import math


class Centered:
    def __init__(self, order=2):
        if order != 2:
            raise ValueError("the MI355X path implements Centered(order=2) as a model advection scheme")
        self.order = order

    def __repr__(self):
        return f"Centered(order={self.order})"


class WENO:
    def __init__(self, order=5):
        if order != 5:
            raise ValueError("the MI355X path implements WENO(order=5)")
        self.order = order

    def __repr__(self):
        return "WENO(order=5)"


class AnisotropicMinimumDissipation:
    def __init__(self, C=1 / 3, Cnu=None, Ckappa=None, Cb=None):
        if Cb is not None:
            raise ValueError("the buoyancy modification Cb is not implemented (reference default: nothing)")
        self.Cnu = C if Cnu is None else Cnu
        self.Ckappa = C if Ckappa is None else Ckappa
        if not isinstance(self.Ckappa, (int, float)):
            raise ValueError("per-tracer or functional Cκ is not supported; give one number")

    def __repr__(self):
        return f"AnisotropicMinimumDissipation(Cν={self.Cnu}, Cκ={self.Ckappa}, Cb=nothing)"


class SmagorinskyLilly:
    def __init__(self, C=0.16, Cb=1.0, Pr=1.0):
        self.C, self.Cb, self.Pr = float(C), float(Cb), float(Pr)

    def __repr__(self):
        return f"SmagorinskyLilly(C={self.C}, Cb={self.Cb}, Pr={self.Pr})"


class ScalarDiffusivity:
    def __init__(self, nu=0.0, kappa=0.0):
        self.nu, self.kappa = float(nu), float(kappa)


class BuoyancyTracer:
    required_tracers = ("b",)


class LinearEquationOfState:
    def __init__(self, thermal_expansion=1.67e-4, haline_contraction=7.80e-4):
        self.thermal_expansion = float(thermal_expansion)
        self.haline_contraction = float(haline_contraction)


class SeawaterBuoyancy:
    required_tracers = ("T", "S")

    def __init__(self, equation_of_state=None, gravitational_acceleration=9.80665):
        self.equation_of_state = equation_of_state or LinearEquationOfState()
        self.gravitational_acceleration = float(gravitational_acceleration)


class FPlane:
    def __init__(self, f=None, rotation_rate=7.292115e-5, latitude=None):
        if (f is None) == (latitude is None):
            raise ValueError("Either both keywords rotation_rate and latitude must be specified, *or* only f must be specified.")
        self.f = float(f) if f is not None else 2 * rotation_rate * math.sin(math.radians(latitude))


class _BC:
    kind = 0

    def __init__(self, condition):
        if not isinstance(condition, (int, float)):
            raise ValueError("only constant boundary conditions are supported on the MI355X path")
        self.condition = float(condition)


class FluxBoundaryCondition(_BC):
    kind = 1


class ValueBoundaryCondition(_BC):
    kind = 2


class GradientBoundaryCondition(_BC):
    kind = 3


class FieldBoundaryConditions:
    def __init__(self, top=None, bottom=None, **others):
        for side, bc in others.items():
            if bc is not None:
                raise ValueError(f"{side} boundary conditions: x and y are Periodic on the MI355X path")
        self.top, self.bottom = top, bottom


# ---------------------------------------------------------------------------------------------
# pressure solvers (src/Solvers/fft_based_poisson_solver.jl, fourier_tridiagonal_poisson_solver.jl,
# conjugate_gradient_poisson_solver.jl)
# ---------------------------------------------------------------------------------------------
class FFTBasedPoissonSolver:
    """The FFT solver of a grid with uniform spacing (used directly, or as a CG preconditioner)."""

    def __init__(self, grid=None):
        self.grid = grid


class FourierTridiagonalPoissonSolver:
    """FFT in x, y and a tridiagonal solve in a stretched z."""

    def __init__(self, grid=None):
        self.grid = grid


class DiagonallyDominantPreconditioner:
    """Pointwise truncated-Neumann-series preconditioner (conjugate_gradient_poisson_solver.jl:256-296)."""


class ColumnwiseTridiagonalPreconditioner:
    """ColumnwiseTridiagonalPreconditioner(grid): per horizontal column, the vertical part of V∇²
    (diagonal regularised by 1 + 1/100, inactive cells as identity rows) solved exactly by one
    Thomas sweep (conjugate_gradient_poisson_solver.jl:298-397)."""

    def __init__(self, grid=None):
        self.grid = grid


class LearnedPreconditioner:
    """The FFT preconditioner of the underlying grid plus a learned convolutional correction
    (a four-level hierarchy of seven 3×3×3 convolutions of 16 channels on the grid and on the grid
    coarsened 2×, 4× and 8×, BF16 on the matrix cores), applied inside flexible CG
    with a fallback to the FFT preconditioner when convergence stalls.  No reference counterpart:
    it takes the place of the preconditioner call of conjugate_gradient_poisson_solver.jl:169-211.
    `weights`: optional (7, 27, 16, 16) FP32 array [layer][tap][cin][cout]; zero weights give the
    FFT preconditioner."""

    def __init__(self, weights=None):
        self.weights = weights


class ConjugateGradientPoissonSolver:
    """ConjugateGradientPoissonSolver(grid; preconditioner, reltol, abstol, maxiter)
    (conjugate_gradient_poisson_solver.jl:134-163).  preconditioner: "default" (the reference's
    choice), None, DiagonallyDominantPreconditioner(), ColumnwiseTridiagonalPreconditioner(), or an FFTBasedPoissonSolver /
    FourierTridiagonalPoissonSolver of the grid.  reltol and abstol default to sqrt(eps(FT)),
    maxiter to Nx·Ny·Nz."""

    def __init__(self, grid=None, preconditioner="default", reltol=None, abstol=None, maxiter=None):
        self.grid = grid
        self.preconditioner = preconditioner
        self.reltol, self.abstol, self.maxiter = reltol, abstol, maxiter

    def preconditioner_kind(self):
        from . import _abi
        p = self.preconditioner
        if isinstance(p, str) and p == "default":
            return _abi.OM_PRECOND_DEFAULT
        if p is None:
            return _abi.OM_PRECOND_NONE
        if isinstance(p, DiagonallyDominantPreconditioner):
            return _abi.OM_PRECOND_DIAGONALLY_DOMINANT
        if isinstance(p, (FFTBasedPoissonSolver, FourierTridiagonalPoissonSolver)):
            return _abi.OM_PRECOND_FFT
        if isinstance(p, ColumnwiseTridiagonalPreconditioner):
            return _abi.OM_PRECOND_COLUMNWISE_TRIDIAGONAL
        if isinstance(p, LearnedPreconditioner):
            return _abi.OM_PRECOND_LEARNED
        raise ValueError(f"unsupported preconditioner {p!r}")

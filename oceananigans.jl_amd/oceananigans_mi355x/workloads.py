"""Synthetic workloads (seeded) for the benchmark, the smoke test and the parity tests.

free_convection   : 3D free-convection LES — WENO5, AnisotropicMinimumDissipation, BuoyancyTracer with
                    b = N²z + U(-1e-3, 1e-3)·N²·Δz noise in the top 16 cells, surface buoyancy flux
                    1e-8 m²/s³, FPlane f = 1e-4 s⁻¹ (the 256×256×128 benchmark case at full size).
wind_convection   : WENO5, SmagorinskyLilly(C=0.16), linear EOS (α=2e-4, β=8e-4), T and S, surface
                    momentum and heat fluxes (examples/ocean_wind_mixing_and_convection.jl:82-123).
turbulence2d      : (Periodic, Periodic, Flat), WENO5, u and v ~ U(-1, 1).
stretched_convection : WENO5 + AMD + BuoyancyTracer (N² = 1e-6 s⁻²) on a geometrically stretched z with
                    u = 0.1 m/s + U(-1e-3, 1e-3) inflow noise.
ensemble_members  : free-convection members of uneven sizes drawn from {64³, 96×96×64, 128×128×64,
                    128³}, each with its own noise seed, N² ~ U(1e-6, 1e-5) s⁻² and surface buoyancy
                    flux ~ U(1e-9, 1e-8) m²/s³ (FP32, WENO5 + AMD).
seamount          : 10 km × 10 km × 1 km, z faces geometrically stretched (ratio 1.02, finest at the
                    surface), GridFittedBottom Gaussian seamount (600 m high, e-folding 1.5 km) at the
                    centre, WENO5 + AMD + BuoyancyTracer (N² = 1e-6 s⁻²), u = 0.1 m/s + 1e-3 noise
                    (seed 3), conjugate-gradient pressure solver to a 1e-8 relative residual.
"""
import numpy as np

from .grids import Flat, GridFittedBottom, ImmersedBoundaryGrid, Periodic, RectilinearGrid
from .models import field_nodes
from .physics import (AnisotropicMinimumDissipation, BuoyancyTracer, FieldBoundaryConditions, FluxBoundaryCondition,
                      FPlane, GradientBoundaryCondition, LinearEquationOfState, SeawaterBuoyancy, SmagorinskyLilly, WENO)
from .synthetic import uniform


def free_convection(N=(256, 256, 128), L=None, float_type=np.float32, noise_velocity=0.0, seed=1, N2=1e-5,
                    buoyancy_flux=1e-8, f=1e-4):
    L = L or tuple(float(n) for n in N)
    grid = RectilinearGrid(float_type=float_type, size=N, x=(0, L[0]), y=(0, L[1]), z=(-L[2], 0))
    kw = dict(advection=WENO(), closure=AnisotropicMinimumDissipation(), buoyancy=BuoyancyTracer(),
              coriolis=FPlane(f=f), tracers=("b",),
              boundary_conditions={"b": FieldBoundaryConditions(top=FluxBoundaryCondition(buoyancy_flux))})
    dz = L[2] / N[2]
    z = grid.znodes("c")
    Z = np.broadcast_to(z[None, None, :], tuple(N))
    b = N2 * Z + uniform(seed, Z.shape, -1e-3, 1e-3) * N2 * dz * (Z > -16 * dz)
    init = {"b": b}
    if noise_velocity:
        init["u"] = uniform(seed + 10, tuple(N), -noise_velocity, noise_velocity)
        init["v"] = uniform(seed + 11, tuple(N), -noise_velocity, noise_velocity)
        init["w"] = uniform(seed + 12, tuple(N), -noise_velocity, noise_velocity)
    return grid, kw, init


ENSEMBLE_SIZES = ((64, 64, 64), (96, 96, 64), (128, 128, 64), (128, 128, 128))


def ensemble_members(n=64, seed=5, float_type=np.float32, sizes=ENSEMBLE_SIZES):
    """[(grid, kw, init)] for n free-convection members; sizes and parameters from one seeded draw."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        N = tuple(sizes[int(rng.integers(len(sizes)))])
        N2 = float(rng.uniform(1e-6, 1e-5))
        flux = float(rng.uniform(1e-9, 1e-8))
        out.append(free_convection(N=N, float_type=float_type, seed=seed * 1000 + q, N2=N2, buoyancy_flux=flux))
    return out


def wind_convection(N=(512, 512, 256), float_type=np.float32, seed=2):
    L = tuple(float(n) for n in N)
    grid = RectilinearGrid(float_type=float_type, size=N, x=(0, L[0]), y=(0, L[1]), z=(-L[2], 0))
    Qh, rho0, cP = 200.0, 1026.0, 3991.0
    bcs = {"u": FieldBoundaryConditions(top=FluxBoundaryCondition(-1e-4)),
           "T": FieldBoundaryConditions(top=FluxBoundaryCondition(Qh / (rho0 * cP)),
                                        bottom=GradientBoundaryCondition(0.01))}
    kw = dict(advection=WENO(), closure=SmagorinskyLilly(C=0.16, Cb=1.0, Pr=1.0),
              buoyancy=SeawaterBuoyancy(equation_of_state=LinearEquationOfState(2e-4, 8e-4)),
              tracers=("T", "S"), boundary_conditions=bcs)
    z = grid.znodes("c")
    Z = np.broadcast_to(z[None, None, :], tuple(N))
    init = {"T": 20 + 0.01 * Z, "S": np.full(Z.shape, 35.0),
            "u": uniform(seed, Z.shape, -1e-3, 1e-3), "v": uniform(seed + 1, Z.shape, -1e-3, 1e-3),
            "w": uniform(seed + 2, Z.shape, -1e-3, 1e-3)}
    return grid, kw, init


def turbulence2d(N=(128, 128), float_type=np.float64, seed=0):
    grid = RectilinearGrid(float_type=float_type, size=N, x=(0, 2 * np.pi), y=(0, 2 * np.pi),
                           topology=(Periodic, Periodic, Flat))
    kw = dict(advection=WENO())
    init = {"u": uniform(seed, (N[0], N[1], 1), -1, 1), "v": uniform(seed + 1, (N[0], N[1], 1), -1, 1)}
    return grid, kw, init


def stretched_convection(N=(16, 16, 16), float_type=np.float64, seed=3, ratio=1.02, Lz=100.0, Lh=32.0):
    Nz = N[2]
    d_top = Lz * (ratio - 1) / (ratio ** Nz - 1)
    depths = np.concatenate([[0.0], np.cumsum(d_top * ratio ** np.arange(Nz))])
    faces = -depths[::-1]
    faces[0] = -Lz
    faces[-1] = 0.0
    grid = RectilinearGrid(float_type=float_type, size=N, x=(0, Lh), y=(0, Lh), z=faces)
    kw = dict(advection=WENO(), closure=AnisotropicMinimumDissipation(), buoyancy=BuoyancyTracer(),
              tracers=("b",), coriolis=FPlane(f=1e-4))
    z = grid.znodes("c")
    Z = np.broadcast_to(z[None, None, :], tuple(N))
    init = {"b": 1e-6 * Z, "u": 0.1 + 1e-3 * uniform(seed + 1, tuple(N), -1, 1)}
    return grid, kw, init


def seamount(N=(256, 256, 128), float_type=np.float32, seed=3, preconditioner="default", reltol=1e-8, maxiter=1000):
    from .physics import ConjugateGradientPoissonSolver
    Lx = Ly = 10e3
    Lz = 1e3
    r = 1.02
    d = [Lz * (r - 1) / (r ** N[2] - 1) * r ** q for q in range(N[2])]   # finest at the surface
    faces = [0.0]
    for dz in d:
        faces.append(faces[-1] - dz)
    z = np.array(faces[::-1])
    z = z - z[0] - Lz
    z[-1] = 0.0
    under = RectilinearGrid(float_type=float_type, size=N, x=(0, Lx), y=(0, Ly), z=z)
    hill = lambda x, y: -Lz + 600 * np.exp(-((x - Lx / 2) ** 2 + (y - Ly / 2) ** 2) / 1500 ** 2)
    grid = ImmersedBoundaryGrid(under, GridFittedBottom(hill))
    kw = dict(advection=WENO(), closure=AnisotropicMinimumDissipation(), buoyancy=BuoyancyTracer(), tracers=("b",),
              pressure_solver=ConjugateGradientPoissonSolver(preconditioner=preconditioner, reltol=reltol, abstol=0,
                                                             maxiter=maxiter))
    init = {"u": 0.1 + 1e-3 * uniform(seed, tuple(N), -1, 1), "b": lambda x, y, z: 1e-6 * z}
    return grid, kw, init

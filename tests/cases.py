"""Model configurations shared by the parity tests, the smoke test and the benchmark.

Each case is the small-size analogue of one workload: a grid, the model keyword arguments and
seeded initial conditions (SplitMix64 streams, identical for the HIP library and the oracle).
"""
import numpy as np

import oceananigans_mi355x as om
from oceananigans_mi355x import workloads
from oceananigans_mi355x.synthetic import uniform  # noqa: F401


def convection(N=(16, 16, 16), L=None, float_type=np.float64, noise_velocity=0.0, seed=1):
    """Free convection LES: WENO5 + AMD + BuoyancyTracer + FPlane + surface buoyancy flux."""
    return workloads.free_convection(N=N, L=L, float_type=float_type, noise_velocity=noise_velocity, seed=seed)


def wind_convection(N=(16, 16, 16), float_type=np.float64, seed=2):
    """Wind-and-convection LES: WENO5 + SmagorinskyLilly + linear EOS with T, S and surface fluxes."""
    return workloads.wind_convection(N=N, float_type=float_type, seed=seed)


def turbulence2d(N=(32, 32), float_type=np.float64, seed=0):
    """2D periodic turbulence on (Periodic, Periodic, Flat): WENO5, no closure, no buoyancy."""
    return workloads.turbulence2d(N=N, float_type=float_type, seed=seed)


def stretched_convection(N=(16, 16, 16), float_type=np.float64, seed=3, ratio=1.02, Lz=100.0):
    """WENO5 + AMD + BuoyancyTracer on a geometrically stretched z (finest at the surface)."""
    return workloads.stretched_convection(N=N, float_type=float_type, seed=seed, ratio=ratio, Lz=Lz)


def seamount(N=(16, 16, 12), float_type=np.float64, stretched=True, seed=3, noise=1e-3, N2=1e-6):
    """The immersed-seamount configuration at small size: stretched z (ratio 1.02, finest at the
    surface), GridFittedBottom Gaussian seamount, WENO5 + AMD + BuoyancyTracer, uniform inflow
    u = 0.1 plus seeded noise, CG pressure solver."""
    Lx = Ly = 10e3
    Lz = 1e3
    Nz = N[2]
    if stretched:
        r = 1.02
        d = [Lz * (r - 1) / (r ** Nz - 1) * r ** q for q in range(Nz)]     # finest at the surface
        faces = [0.0]
        for dz in d:
            faces.append(faces[-1] - dz)
        z = np.array(faces[::-1])
        z = z - z[0] - Lz
        z[-1] = 0.0
    else:
        z = (-Lz, 0.0)
    under = om.RectilinearGrid(float_type=float_type, size=N, x=(0, Lx), y=(0, Ly), z=z)
    hill = lambda x, y: -Lz + 600 * np.exp(-((x - Lx / 2) ** 2 + (y - Ly / 2) ** 2) / 1500 ** 2)
    grid = om.ImmersedBoundaryGrid(under, om.GridFittedBottom(hill))
    kw = dict(advection=om.WENO(), closure=om.AnisotropicMinimumDissipation(), buoyancy=om.BuoyancyTracer(),
              tracers=("b",), pressure_solver=om.ConjugateGradientPoissonSolver(reltol=1e-13, abstol=0))
    from oceananigans_mi355x.synthetic import uniform
    shape_u = under.N
    init = {"u": 0.1 + noise * uniform(seed, shape_u, -1, 1), "b": lambda x, y, z: N2 * z}
    return grid, kw, init


def taylor_green(N=16, float_type=np.float64):
    """Taylor–Green vortex with ν = 1 (test/test_dynamics.jl:214-259): Centered(2), ScalarDiffusivity."""
    grid = om.RectilinearGrid(float_type=float_type, size=(N, N, 2), extent=(1, 1, 1))
    kw = dict(closure=om.ScalarDiffusivity(nu=1.0))
    init = {"u": lambda x, y, z: -np.sin(2 * np.pi * y), "v": lambda x, y, z: np.sin(2 * np.pi * x)}
    return grid, kw, init


CASES = {"convection": convection, "wind_convection": wind_convection, "turbulence2d": turbulence2d,
         "stretched_convection": stretched_convection}


def prognostic_names(kw):
    return ("u", "v", "w") + tuple(kw.get("tracers", ()))


def rel_err(a, b):
    """max|a − b| / max|b| (absolute when b ≡ 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    d = np.max(np.abs(a - b)) if a.size else 0.0
    return d / scale if scale > 0 else d


def oracle_for(grid, kw, init, oracle_float_type=np.float64, round_inputs_to=None):
    """Build the oracle model for the same description and apply set!(model; init...)."""
    from oracle.oracle import OracleModel

    g2 = grid
    if oracle_float_type != grid.float_type:
        g2 = om.RectilinearGrid(float_type=oracle_float_type, size=tuple(n for n, t in zip(grid.N, grid.topology) if t is not om.Flat),
                                x=(grid.origin[0], grid.origin[0] + grid.extent[0]),
                                y=(grid.origin[1], grid.origin[1] + grid.extent[1]),
                                z=grid.z_faces if grid.z_faces is not None else (grid.origin[2], grid.origin[2] + grid.extent[2]),
                                topology=grid.topology, halo=tuple(h for h, t in zip(grid.H, grid.topology) if t is not om.Flat))
        if isinstance(grid, om.ImmersedBoundaryGrid):
            g2 = om.ImmersedBoundaryGrid(g2, grid.immersed_boundary)
    cfg, keep = om.build_config(g2, **kw)
    m = OracleModel(cfg, g2)
    m._keep = keep
    names = prognostic_names(kw)
    for name, val in init.items():
        v = om.models._evaluate(g2, name, val)
        if round_inputs_to is not None:
            v = np.asarray(v, dtype=round_inputs_to).astype(np.float64)
        m.set(names.index(name), v)
    m.set_finalize(True)
    return m


def product_for(grid, kw, init):
    model = om.NonhydrostaticModel(grid, **kw)
    om.set_(model, **init)
    return model

"""Near-boundary behaviour of the HIP kernels through the C ABI: Value boundary conditions against
the oracle, and the z reach of the tendency kernels next to a Bounded wall.

Reference anchors: src/BoundaryConditions/fill_halo_regions_value_gradient.jl:7-60,
src/Advection/topologically_conditional_interpolation.jl:44-58,106-124.
"""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import convection, oracle_for, product_for, prognostic_names, rel_err, wind_convection
from oceananigans_mi355x.synthetic import uniform

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def value_bc_case(N=(16, 16, 16), float_type=np.float64):
    """Wind-and-convection physics (SmagorinskyLilly, linear EOS) with Value conditions: T fixed at
    the bottom, S fixed at the top, no-slip v at the bottom (v = 0 there); O(1)-relative fields."""
    grid, kw, _ = wind_convection(N=N, float_type=float_type)
    bcs = dict(kw["boundary_conditions"])
    bcs["T"] = om.FieldBoundaryConditions(top=bcs["T"].top, bottom=om.ValueBoundaryCondition(2e-3))
    bcs["S"] = om.FieldBoundaryConditions(top=om.ValueBoundaryCondition(-1e-3))
    bcs["v"] = om.FieldBoundaryConditions(bottom=om.ValueBoundaryCondition(0.0))
    kw = dict(kw, boundary_conditions=bcs)
    init = dict(T=1e-3 * uniform(30, grid.N, -1, 1), S=1e-3 * uniform(31, grid.N, -1, 1),
                u=5e-2 * uniform(32, grid.N, -1, 1), v=5e-2 * uniform(33, grid.N, -1, 1),
                w=5e-2 * uniform(34, grid.N, -1, 1))
    return grid, kw, init


def test_value_bc_halos_and_tendencies_fp64():
    """After set!: the Value halos (2v − c₁), the projected state, the tendencies and νₑ equal the
    oracle's (FP64 ≤ 1e-12)."""
    grid, kw, init = value_bc_case()
    model = product_for(grid, kw, init)
    orc = oracle_for(grid, kw, init)
    names = prognostic_names(kw)
    H = grid.H
    for q, n in enumerate(names):
        f = model.velocities[n] if q < 3 else model.tracers[n]
        par = f.parent()
        lo = [1 - H[0], 1 - H[1], 0]
        hi = [grid.Nx + H[0], grid.Ny + H[1], grid.Nz + 1 + (1 if n == "w" else 0)]
        ref = orc.get_range(q, lo, hi)
        e = rel_err(par[:, :, H[2] - 1: H[2] - 1 + ref.shape[2]], ref)
        assert e <= FP64_TOL, (n, e)
        G = model.Gn[n].interior()
        Go = orc.get(om._abi.OM_FIELD_GN + q)
        assert rel_err(G, Go) <= FP64_TOL, (n, rel_err(G, Go))
    assert rel_err(model.closure_fields.nu.interior(), orc.get(om._abi.OM_FIELD_NU)) <= FP64_TOL


def test_value_bc_one_step_fp64():
    grid, kw, init = value_bc_case()
    model = product_for(grid, kw, init)
    orc = oracle_for(grid, kw, init)
    om.time_step_(model, 1.0)
    orc.time_step(1.0, 1)
    names = prognostic_names(kw)
    errs = {n: rel_err((model.velocities[n] if q < 3 else model.tracers[n]).interior(), orc.get(q))
            for q, n in enumerate(names)}
    errs["pNHS"] = rel_err(model.pressures.pNHS.interior(), orc.get(om._abi.OM_FIELD_PNHS))
    print("Value BCs, one step", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(e <= FP64_TOL for e in errs.values()), errs


def _nan_layers(field, grid, layers, face=False):
    par = field.parent()
    Hz, Nz = grid.H[2], grid.Nz
    for L in layers:
        par[:, :, Hz - L] = np.nan
        par[:, :, Hz + Nz + (1 if face else 0) + L - 1] = np.nan
    field.set_parent(par)


@pytest.mark.parametrize("float_type", [np.float64, np.float32])
@pytest.mark.parametrize("kernel", ["tiled", "generic"])
@pytest.mark.parametrize("case", ["convection", "value_bcs"])
def test_near_wall_tendencies_read_one_halo_layer(case, kernel, float_type):
    """NaN in z-halo layers 2 and 3 of every prognostic and closure field, and in all of pHY′'s z
    halos, leaves the tendencies of both kernels bit for bit unchanged."""
    if case == "convection":
        grid, kw, init = convection(N=(20, 18, 14), noise_velocity=1e-2, float_type=float_type)
    else:
        grid, kw, init = value_bc_case(N=(20, 18, 14), float_type=float_type)
    model = product_for(grid, kw, init)
    model.set_tuning(tendency_kernel=kernel)
    model.update_state()
    names = prognostic_names(kw)
    G0 = {n: model.Gn[n].interior() for n in names}
    assert all(np.isfinite(g).all() for g in G0.values())
    for q, n in enumerate(names):
        _nan_layers(model.velocities[n] if q < 3 else model.tracers[n], grid, (2, 3), face=(n == "w"))
    for f in model.closure_fields.values():
        _nan_layers(f, grid, (2, 3))
    if kw.get("buoyancy") is not None:
        _nan_layers(model.pressures.pHY, grid, (1, 2, 3))
    model.compute_tendencies()
    for n in names:
        assert np.array_equal(model.Gn[n].interior(), G0[n]), f"{case} {kernel}: G{n} changed"
    model.close()


def _check_pressure_halos(model, label):
    """pNHS's halos are the periodic images (x, y) and the zero-gradient copies (the first layer
    below and above a Bounded z) of its interior, bit for bit."""
    p = model.pressures.pNHS
    info = p.info()
    hx, hy, hz = (1 - int(info.offset[d]) for d in range(3))   # offset: first parent index, 1 − H
    inner = p.interior()
    par = p.parent()
    nx, ny, nz = inner.shape
    zs = slice(hz, hz + nz)
    for h in range(hx):
        assert np.array_equal(par[h, hy:hy + ny, zs], inner[nx - hx + h]), (label, "west", h)
        assert np.array_equal(par[hx + nx + h, hy:hy + ny, zs], inner[h]), (label, "east", h)
    for h in range(hy):
        assert np.array_equal(par[hx:hx + nx, h, zs], inner[:, ny - hy + h]), (label, "south", h)
        assert np.array_equal(par[hx:hx + nx, hy + ny + h, zs], inner[:, h]), (label, "north", h)
    assert np.array_equal(par[hx:hx + nx, hy:hy + ny, hz - 1], inner[:, :, 0]), (label, "bottom")
    assert np.array_equal(par[hx:hx + nx, hy:hy + ny, hz + nz], inner[:, :, nz - 1]), (label, "top")


@pytest.mark.parametrize("float_type", [np.float64, np.float32])
def test_pressure_halos_filled_after_the_correction_reads_wrapped_neighbours(float_type):
    """With periodic x and y the pressure correction reads p's neighbours across the edges by
    wrapped index and p's halo fill joins the next halo launch.  The halos must still hold the
    filled values when a call returns: after time_step! (the fill joins update_state!'s) and after
    a bare pressure correction (filled at the end of the call), as the reference's
    fill_halo_regions!(pNHS) after the solve leaves them (pressure_correction.jl:6-20)."""
    grid, kw, init = convection(N=(16, 16, 16), noise_velocity=1e-2, float_type=float_type)
    model = product_for(grid, kw, init)
    om.time_step_(model, 1.0)
    assert np.abs(model.pressures.pNHS.interior()).max() > 0
    _check_pressure_halos(model, "time_step")
    model.compute_pressure_correction(0.5)
    _check_pressure_halos(model, "pressure_correction")
    model.close()

"""GridFittedBottom immersed boundary (src/ImmersedBoundaries/grid_fitted_bottom.jl) on the GPU
against the oracle, through the C ABI.

Covered: the conditional differences and fluxes next to the bottom, the reduced interpolation
order near immersed cells, the masked Coriolis transports, the masking of immersed tracers and
velocities, the conditional Laplacian of the conjugate-gradient solver with the FFT preconditioner
of the underlying grid (the reference's default on an immersed grid), and the FFT and
Fourier–tridiagonal solvers applied to the immersed divergence.

Bars: FP64 tendencies and closure fields on identical inputs ≤ 1e-12 (the FP64 kernels evaluate
every term in the oracle's order without FMA contraction); the state and pressure after steps
≤ 1e-11, as for the other CG and offset cases (u = 0.1 + 1e-3 noise conditions the WENO weights,
and the CG sums run in a different order)."""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import oracle_for, product_for, prognostic_names, rel_err, seamount
from test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu

CASES = {
    "seamount_stretched": lambda: seamount(N=(16, 16, 12)),
    "seamount_uniform_ragged": lambda: seamount(N=(18, 13, 10), stretched=False),
}


def _immersed(grid, orc):
    dzc, dzf, zc, zf = orc.grid_z()
    H = grid.H[2]
    zcent = zc[H:H + grid.Nz]
    ft = grid.float_type
    return zcent.astype(ft)[None, None, :] <= grid.bottom_heights().astype(ft)[:, :, None]


@pytest.mark.parametrize("case", sorted(CASES))
def test_immersed_tendencies_identical_inputs_fp64(case):
    grid, kw, init = CASES[case]()
    model = om.NonhydrostaticModel(grid, **kw)
    om.set_(model, enforce_incompressibility=False, **init)
    from oracle.oracle import OracleModel
    cfg, keep = om.build_config(grid, **kw)
    orc = OracleModel(cfg, grid)
    names = prognostic_names(kw)
    for name, val in init.items():
        orc.set(names.index(name), om.models._evaluate(grid, name, val))
    orc.set_finalize(False)
    imm = _immersed(grid, orc)
    print(case, "immersed cells", int(imm.sum()), "of", imm.size, flush=True)
    assert imm.any() and not imm.all()
    compare_state(model, orc, kw, fields=("state", "G"), label=f"{case} identical inputs")
    assert rel_err(model.closure_fields.nu.interior(), orc.get(om._abi.OM_FIELD_NU)) <= 1e-12
    assert rel_err(model.closure_fields.kappa_b.interior(), orc.get(om._abi.OM_FIELD_KAPPA0)) <= 1e-12
    assert rel_err(model.pressures.pHY.interior(), orc.get(om._abi.OM_FIELD_PHY)) <= 1e-12


def test_immersed_tiled_kernels_with_generic_fixup_fp64():
    """On an immersed grid the tiled tendency and closure kernels cover every cell and the generic
    kernels then redo the cells within reach of the bottom (an immersed cell within 5 cells in x, y
    and z).  Out of that reach every immersed-boundary predicate is false and both kernels evaluate
    the same expressions.  On a grid wide enough that most cells are out of reach, the tendencies
    and closure fields agree with the generic kernels' and with the oracle's to 1e-12, also after
    the autotuner has re-run the tiled kernels."""
    grid, kw, init = seamount(N=(64, 48, 24))
    models = {}
    for kern in ("tiled", "generic"):
        m = om.NonhydrostaticModel(grid, **kw)
        m.set_tuning(tendency_kernel=kern)
        om.set_(m, enforce_incompressibility=False, **init)
        models[kern] = m
    from oracle.oracle import OracleModel
    cfg, keep = om.build_config(grid, **kw)
    orc = OracleModel(cfg, grid)
    names = prognostic_names(kw)
    for name, val in init.items():
        orc.set(names.index(name), om.models._evaluate(grid, name, val))
    orc.set_finalize(False)
    # cells the generic kernels redo: below the window maximum of the bottom index plus the reach
    kb = _immersed(grid, orc).sum(axis=2)
    R = 5
    win = np.zeros_like(kb)
    for di in range(-R, R + 1):
        for dj in range(-R, R + 1):
            win = np.maximum(win, np.roll(np.roll(kb, di, axis=0), dj, axis=1))
    kfix = np.where(win > 0, np.minimum(win + R, grid.Nz), 0)
    frac = float(np.minimum(kfix, grid.Nz).sum()) / kb.size / grid.Nz
    print("cells redone by the generic kernels:", f"{frac:.2f}", flush=True)
    assert 0.05 < frac < 0.6
    t, g = models["tiled"], models["generic"]
    assert t.get_tuning()["tendency_kernel"] == "tiled" and g.get_tuning()["tendency_kernel"] == "generic"
    compare_state(t, orc, kw, fields=("state", "G"), label="tiled + fix-up vs oracle")
    compare_state(g, orc, kw, fields=("state", "G"), label="generic vs oracle")
    for n in names:
        assert rel_err(t.Gn[n].interior(), g.Gn[n].interior()) <= 1e-12, n
    assert rel_err(t.closure_fields.nu.interior(), orc.get(om._abi.OM_FIELD_NU)) <= 1e-12
    assert rel_err(t.closure_fields.kappa_b.interior(), orc.get(om._abi.OM_FIELD_KAPPA0)) <= 1e-12
    before = {n: t.Gn[n].interior().copy() for n in names}
    t.autotune()
    for n in names:
        assert np.array_equal(t.Gn[n].interior(), before[n]), n


def test_immersed_tiled_kernels_with_generic_fixup_fp32():
    """The same split in FP32: the tiled kernels' FP32 paths (fused multiply-adds, hardware
    reciprocals) round differently from the generic kernel's, so the two agree to FP32 rounding."""
    grid, kw, init = seamount(N=(64, 48, 24), float_type=np.float32)
    models = {}
    for kern in ("tiled", "generic"):
        m = om.NonhydrostaticModel(grid, **kw)
        m.set_tuning(tendency_kernel=kern)
        om.set_(m, enforce_incompressibility=False, **init)
        models[kern] = m
    t, g = models["tiled"], models["generic"]
    for n in prognostic_names(kw):
        e = rel_err(t.Gn[n].interior(), g.Gn[n].interior())
        print("G" + n, f"{e:.2e}", flush=True)
        assert e <= 1e-5, n
    assert rel_err(t.closure_fields.nu.interior(), g.closure_fields.nu.interior()) <= 1e-5
    assert rel_err(t.closure_fields.kappa_b.interior(), g.closure_fields.kappa_b.interior()) <= 1e-5


@pytest.mark.parametrize("case", sorted(CASES))
def test_immersed_steps_match_oracle_fp64(case):
    """Three RK3 steps with the CG solver (FFT preconditioner of the underlying grid, reltol 1e-13)."""
    grid, kw, init = CASES[case]()
    model = product_for(grid, kw, init)
    orc = oracle_for(grid, kw, init)
    om.time_step_(model, 20.0)
    orc.time_step(20.0, 1)
    compare_state(model, orc, kw, fields=("state", "p"), tol=1e-11, label=f"{case} step 1")
    print("iterations gpu/oracle", model.last_solve_info(), orc.last_solve_info(), flush=True)
    it, it_o = model.last_solve_info()[0], orc.last_solve_info()[0]
    assert abs(it - it_o) <= max(1, 0.05 * it_o)
    for _ in range(2):
        om.time_step_(model, 20.0)
    orc.time_step(20.0, 2)
    compare_state(model, orc, kw, fields=("state", "p"), tol=1e-11, label=f"{case} step 3")


@pytest.mark.parametrize("solver", ["fft", "ft"])
def test_immersed_direct_solver_step_fp64(solver):
    """The reference also runs an immersed grid with the (inexact there) FFT-based solver of the
    underlying grid: one step with FFTBasedPoissonSolver (uniform z) or the Fourier–tridiagonal
    solver (stretched z) against the oracle."""
    stretched = solver == "ft"
    grid, kw, init = seamount(N=(16, 12, 10), stretched=stretched)
    kw = dict(kw, pressure_solver=om.FourierTridiagonalPoissonSolver() if stretched else om.FFTBasedPoissonSolver())
    model = product_for(grid, kw, init)
    orc = oracle_for(grid, kw, init)
    om.time_step_(model, 20.0)
    orc.time_step(20.0, 1)
    compare_state(model, orc, kw, fields=("state", "p"), tol=1e-11, label=f"{solver} immersed step")


@pytest.mark.parametrize("float_type", [np.float64, np.float32])
def test_immersed_resting_stratification_stays_at_rest(float_type):
    """b = N²z at rest over the seamount: no spurious flow and b unchanged, exactly."""
    grid, kw, _ = seamount(N=(16, 16, 12), float_type=float_type)
    model = product_for(grid, kw, {"b": lambda x, y, z: 1e-6 * z})
    b0 = model.tracers.b.interior().copy()
    for _ in range(3):
        om.time_step_(model, 30.0)
    for n in ("u", "v", "w"):
        assert np.max(np.abs(model.velocities[n].interior())) == 0
    np.testing.assert_array_equal(model.tracers.b.interior(), b0)


def test_immersed_projection_masks_and_conservation():
    """After steps: u, v, w vanish on the immersed periphery, the active cells are divergence-free,
    and the fluid's Σ V b is conserved (no flux through immersed or wall faces)."""
    grid, kw, init = seamount(N=(16, 16, 12))
    init = dict(init, b=lambda x, y, z: 1e-6 * z + 1e-8 * np.sin(x / 700.0) * np.cos(y / 900.0))
    model = product_for(grid, kw, init)
    from oracle.oracle import OracleModel
    o = OracleModel(om.build_config(grid, **kw)[0], grid)
    imm = _immersed(grid, o)
    dzc = o.grid_z()[0][grid.H[2]:grid.H[2] + grid.Nz]
    V = np.broadcast_to(dzc[None, None, :], grid.N)
    total0 = (V * model.tracers.b.interior())[~imm].sum()
    for _ in range(3):
        om.time_step_(model, 20.0)
    u, v, w = (model.velocities[n].interior() for n in ("u", "v", "w"))
    assert np.all(u[imm | np.roll(imm, 1, axis=0)] == 0)
    assert np.all(v[imm | np.roll(imm, 1, axis=1)] == 0)
    imm_w = np.zeros(w.shape, dtype=bool)
    imm_w[:, :, 1:grid.Nz] = imm[:, :, 1:] | imm[:, :, :-1]
    assert np.all(w[imm_w] == 0)
    div = model.divergence()
    dmax = np.abs(div[~imm]).max()
    print("max |div| over active cells", dmax, flush=True)
    assert dmax <= 1e-10 * np.abs(u).max() / (grid.extent[0] / grid.Nx)
    total = (V * model.tracers.b.interior())[~imm].sum()
    assert abs(total - total0) <= 1e-12 * abs(total0)


@pytest.mark.parametrize("pkind", ["fft", "dd", "ct"])
def test_immersed_cg_solve_matches_oracle(pkind):
    """The conditional V∇² on the stretched seamount with each preconditioner: the GPU solution
    against the oracle's, and the recomputed FP64 residual at the requested tolerance."""
    from oceananigans_mi355x.synthetic import uniform
    grid, kw, _ = seamount(N=(16, 16, 12))
    pc = {"fft": om.FourierTridiagonalPoissonSolver(), "dd": om.DiagonallyDominantPreconditioner(),
          "ct": om.ColumnwiseTridiagonalPreconditioner()}[pkind]
    kw = dict(kw, pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=pc, reltol=1e-13, abstol=0))
    model = om.NonhydrostaticModel(grid, **kw)
    orc = oracle_for(grid, kw, {})
    imm = _immersed(grid, orc)
    rhs = uniform(21, grid.N, -1, 1)
    rhs[imm] = 0
    dz = np.diff(grid.z_faces)
    V = np.broadcast_to(dz[None, None, :], grid.N)
    rhs[~imm] -= (V * rhs)[~imm].sum() / V[~imm].sum()
    phi = model.poisson_solve(rhs)
    it, res, true_res = model.last_solve_info()
    phi_o = orc.poisson_solve(rhs)
    it_o = orc.last_solve_info()[0]
    e = rel_err(phi, phi_o)
    print(pkind, "iterations gpu/oracle", it, it_o, "residual", res, true_res, "rel err", e, flush=True)
    assert np.all(phi[imm] == 0)
    assert e <= 1e-11
    assert abs(it - it_o) <= max(1, 0.05 * it_o)
    assert true_res <= 1e-11 * np.linalg.norm(rhs)

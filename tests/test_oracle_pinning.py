"""Pins the CPU oracle against what the reference holds (doctest vectors, its own tests' invariants)
and against what the mathematics fixes (closed forms, scipy routines, analytic solutions).

Reference anchors:
  src/Advection/reconstruction_coefficients.jl:93-97,119-133 (coefficient doctests)
  test/test_weno_smoothness.jl:5-64, test/test_batched_tridiagonal_solver.jl:7-41,
  test/test_poisson_solvers.jl:68-116, test/test_poisson_solvers_stretched_grids.jl:26-49,
  test/test_dynamics.jl:214-352, test/test_time_stepping.jl:138-171, test/test_halo_regions.jl:22-41
"""
import numpy as np
import pytest
import scipy.fft
import scipy.linalg

import oceananigans_mi355x as om
from cases import convection, oracle_for, prognostic_names, rel_err, taylor_green, wind_convection
from oracle import oracle as O


# ---------------------------------------------------------------------------------------------
# reconstruction coefficients and WENO
# ---------------------------------------------------------------------------------------------
def test_weno5_optimal_combination_matches_reference_doctest():
    """Σ_r C★_r · c_r is the 5th-order left-biased stencil; in FP32 it must equal the doctest
    (-0.05f0, 0.45f0, 0.78333336f0, -0.21666667f0, 0.033333335f0) on (ψ[i+1], ψ[i], ψ[i-1], ψ[i-2], ψ[i-3])."""
    w5, _, _, _ = O.coefficients(fp32=True)
    Cs = np.array([3 / 10, 3 / 5, 1 / 10], dtype=np.float32)
    # stencil r (left bias at face i) covers ψ[i-1-r .. i+1-r]; coefficient index q ↦ offset q - 1 - r
    combo = {}
    for r in range(3):
        for q in range(3):
            off = q - 1 - r
            combo[off] = combo.get(off, 0.0) + float(Cs[r]) * w5[r, q]
    got = np.array([combo[1], combo[0], combo[-1], combo[-2], combo[-3]], dtype=np.float32)
    doctest = np.array([-0.05, 0.45, 0.78333336, -0.21666667, 0.033333335], dtype=np.float32)
    np.testing.assert_allclose(got, doctest, atol=2e-7, rtol=0)


def test_centered4_coefficients_match_reference_doctest_bitwise():
    """calc_reconstruction_stencil(Float32, 2, :symmetric, :x) prints
    -0.083333254f0 ψ[i-2] + 0.5833333f0 ψ[i-1] + 0.5833333f0 ψ[i] + -0.083333336f0 ψ[i+1]."""
    _, _, c4, c2 = O.coefficients(fp32=True)
    expect = np.array([-0.083333254, 0.5833333, 0.5833333, -0.083333336], dtype=np.float32)
    assert np.array_equal(c4.astype(np.float32), expect)
    assert np.array_equal(c2, [0.5, 0.5])


@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("fp32", [False, True])
def test_weno5_exact_for_quadratic_cell_averages(left, fp32):
    """Every 3-point candidate reconstructs cell averages of a quadratic exactly, so any convex
    weighting does: cell averages of q(x) = (x - 0.3)² on unit cells give q at the face."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        a, b, c = rng.uniform(-2, 2, 3)
        i = np.arange(-3, 3)                  # cells i-3 .. i+2 around face i (= x = -0.5 for i = 0)
        avg = a * (i ** 2 + 1 / 12) + b * i + c
        face = a * 0.25 - b * 0.5 + c         # q(-0.5)
        val = O.weno5(avg, left=left, fp32=fp32)
        assert abs(val - face) <= (2e-6 if fp32 else 1e-13) * max(1.0, abs(face))


def test_weno3_exact_for_linear_data():
    for left in (True, False):
        S = np.array([1.0, 3.0, 5.0, 7.0])     # face between S[1] and S[2] → 4
        assert abs(O.weno3(S, left=left) - 4.0) < 1e-14


def test_weno5_smoothness_is_three_times_jiang_shu():
    """β_r = 3 · (13/12 (a − 2b + c)² + 1/4 (…)²), the published Jiang–Shu indicator ×3
    (weno_interpolants.jl:172-174 carries un-normalised quadratic forms)."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        S = rng.normal(size=6)
        beta, _ = O.weno5_smoothness(S, left=True)
        # left-biased stencils at face i: r0 = (ψ[i-1], ψ[i], ψ[i+1]), r1 = (ψ[i-2..i]), r2 = (ψ[i-3..i-1])
        f = S  # f[0] = ψ[i-3]
        js0 = 13 / 12 * (f[2] - 2 * f[3] + f[4]) ** 2 + 1 / 4 * (3 * f[2] - 4 * f[3] + f[4]) ** 2
        js1 = 13 / 12 * (f[1] - 2 * f[2] + f[3]) ** 2 + 1 / 4 * (f[1] - f[3]) ** 2
        js2 = 13 / 12 * (f[0] - 2 * f[1] + f[2]) ** 2 + 1 / 4 * (f[0] - 4 * f[1] + 3 * f[2]) ** 2
        np.testing.assert_allclose(beta, 3 * np.array([js0, js1, js2]), rtol=1e-12, atol=1e-13)


def test_weno_smoothness_fp32_matches_reference_test():
    """test/test_weno_smoothness.jl:5-64 (order 5): ρθ = 300 + 0.1 sin(2π(i−1)/6)."""
    S = 300.0 + 0.1 * np.sin(2 * np.pi * np.arange(6) / 6)
    b64, w64 = O.weno5_smoothness(S, left=True, fp32=False)
    b32, w32 = O.weno5_smoothness(np.float32(S).astype(np.float64), left=True, fp32=True)
    assert np.all(b32 >= 0)
    for r in range(3):
        if b64[r] > 0:
            assert abs(b32[r] - b64[r]) <= 1e-2 * abs(b64[r])
    assert abs(w64.sum() - 1) < 1e-14 and abs(w32.sum() - 1) < 1e-6
    np.testing.assert_allclose(w32, w64, atol=1e-3)


def test_weno_weights_optimal_for_linear_data():
    """Linear data: all β equal, τ = 0, so ω = C★ = (3/10, 3/5, 1/10)."""
    _, w = O.weno5_smoothness(np.arange(6, dtype=float) * 0.7 + 2, left=True)
    np.testing.assert_allclose(w, [0.3, 0.6, 0.1], rtol=1e-14)


# ---------------------------------------------------------------------------------------------
# Poisson solvers
# ---------------------------------------------------------------------------------------------
def _laplacian(phi, dx, dy, dzc, dzf, tz):
    """7-point discrete Laplacian: periodic in x, y; Neumann (mirror) or periodic or absent in z.
    dzc[k], dzf[k] for interior k (dzf[0] unused)."""
    L = (np.roll(phi, -1, 0) - 2 * phi + np.roll(phi, 1, 0)) / dx ** 2
    L += (np.roll(phi, -1, 1) - 2 * phi + np.roll(phi, 1, 1)) / dy ** 2
    if tz == "F":
        return L
    if tz == "P":
        return L + (np.roll(phi, -1, 2) - 2 * phi + np.roll(phi, 1, 2)) / dzc[None, None, :] ** 2
    up = np.concatenate([phi[:, :, 1:], phi[:, :, -1:]], axis=2)
    dn = np.concatenate([phi[:, :, :1], phi[:, :, :-1]], axis=2)
    Nz = phi.shape[2]
    fu = np.zeros(Nz)
    fd = np.zeros(Nz)
    fu[:-1] = 1 / dzf[1:]
    fd[1:] = 1 / dzf[1:]
    return L + ((up - phi) * fu[None, None, :] - (phi - dn) * fd[None, None, :]) / dzc[None, None, :]


@pytest.mark.parametrize("tz", ["B", "P", "F"])
def test_fft_poisson_residual_and_scipy(tz):
    """∇²ϕ = b to rounding (dependencies_for_poisson_solvers.jl:115-133) and agreement with an
    independent scipy spectral solve (DCT-II along Bounded z, FFT along periodic axes)."""
    topo = {"B": om.Bounded, "P": om.Periodic, "F": om.Flat}[tz]
    size = (12, 10) if tz == "F" else (12, 10, 8)
    grid = om.RectilinearGrid(size=size, x=(0, 1.2), y=(0, 2.0), z=(-0.8, 0), topology=(om.Periodic, om.Periodic, topo))
    orc = oracle_for(grid, dict(advection=om.WENO()), {})
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, (grid.Nx, grid.Ny, grid.Nz))
    b -= b.mean()
    phi = orc.poisson_solve(b)
    dx, dy, dz = 1.2 / 12, 2.0 / 10, 0.8 / grid.Nz
    dzc = np.full(grid.Nz, dz)
    L = _laplacian(phi, dx, dy, dzc, dzc, tz)
    assert np.max(np.abs(L - b)) < 1e-11
    assert abs(phi.mean()) < 1e-13
    # scipy: eigen-decomposition with poisson_eigenvalues.jl:8-23
    lx = (2 * np.sin(np.arange(12) * np.pi / 12) / dx) ** 2
    ly = (2 * np.sin(np.arange(10) * np.pi / 10) / dy) ** 2
    B = scipy.fft.fft(scipy.fft.fft(b, axis=0), axis=1)
    if tz == "B":
        B = scipy.fft.dct(B.real, type=2, axis=2) + 1j * scipy.fft.dct(B.imag, type=2, axis=2)
        lz = (2 * np.sin(np.arange(grid.Nz) * np.pi / (2 * grid.Nz)) / dz) ** 2
    elif tz == "P":
        B = scipy.fft.fft(B, axis=2)
        lz = (2 * np.sin(np.arange(grid.Nz) * np.pi / grid.Nz) / dz) ** 2
    else:
        lz = np.zeros(1)
    den = lx[:, None, None] + ly[None, :, None] + lz[None, None, :]
    den[0, 0, 0] = 1
    P = -B / den
    P[0, 0, 0] = 0
    if tz == "B":
        P = scipy.fft.idct(P.real, type=2, axis=2) + 1j * scipy.fft.idct(P.imag, type=2, axis=2)
    elif tz == "P":
        P = scipy.fft.ifft(P, axis=2)
    ref = scipy.fft.ifft(scipy.fft.ifft(P, axis=1), axis=0).real
    assert rel_err(phi, ref) < 1e-12


def test_fft_poisson_analytic_mode_second_order():
    """Analytic cosine mode (dependencies_for_poisson_solvers.jl:139-176): L¹ error falls with rate ≈ 2."""
    errs = []
    for N in (32, 64):
        grid = om.RectilinearGrid(size=(N, N, N), x=(0, 2 * np.pi), y=(0, 2 * np.pi), z=(0, np.pi))
        orc = oracle_for(grid, dict(advection=om.WENO()), {})
        x, y, z = grid.xnodes("c"), grid.ynodes("c"), grid.znodes("c")
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        psi = np.cos(2 * X) * np.cos(3 * Y) * np.cos(4 * Z)
        rhs = -(4 + 9 + 16) * psi
        phi = orc.poisson_solve(rhs)
        errs.append(np.mean(np.abs(phi - psi)))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - 2) < 0.05


def test_stretched_grid_metrics_follow_grid_generation():
    """Stretched faces [1,2,4,7,11,16,22,29,37] (test_poisson_solvers_stretched_grids.jl:29-30):
    Δzᶜ = faces difference, Δzᶠ = centre difference, halo spacings replicate the end cells."""
    faces = np.array([1, 2, 4, 7, 11, 16, 22, 29, 37], dtype=float)
    grid = om.RectilinearGrid(size=(4, 4, 8), x=(0, 1), y=(0, 1), z=faces, halo=(3, 3, 3))
    orc = oracle_for(grid, dict(advection=om.WENO()), {})
    dzc, dzf, zc, zf = orc.grid_z()
    H = 3
    d = np.diff(faces)
    full_faces = np.concatenate([faces[0] - d[0] * np.arange(H, 0, -1), faces, faces[-1] + d[-1] * np.arange(1, H + 1)])
    centres = 0.5 * (full_faces[1:] + full_faces[:-1])
    np.testing.assert_allclose(dzc, np.diff(full_faces), rtol=1e-15)
    np.testing.assert_allclose(zc, centres, rtol=1e-15)
    np.testing.assert_allclose(dzf[1:], np.diff(centres), rtol=1e-15)


def test_fourier_tridiagonal_residual_on_stretched_faces():
    """∇²ϕ ≈ R on the stretched grid (test_poisson_solvers_stretched_grids.jl:26-49)."""
    faces = np.array([1, 2, 4, 7, 11, 16, 22, 29, 37], dtype=float)
    grid = om.RectilinearGrid(size=(8, 6, 8), x=(0, 1), y=(0, 1), z=faces)
    orc = oracle_for(grid, dict(advection=om.WENO()), {})
    dzc, dzf, _, _ = orc.grid_z()
    H = 3
    dzc_i, dzf_i = dzc[H:H + 8], dzf[H:H + 8]
    rng = np.random.default_rng(5)
    R = rng.uniform(-1, 1, (8, 6, 8))
    R -= (R * dzc_i[None, None, :]).sum() / (dzc_i.sum() * 48)      # volume-mean-free source
    phi = orc.poisson_solve(R)
    L = _laplacian(phi, 1 / 8, 1 / 6, dzc_i, dzf_i, "B")
    assert np.max(np.abs(L - R)) < 1e-10


def test_thomas_matches_scipy_banded():
    """test_batched_tridiagonal_solver.jl:7-41 with b = 3 + rand."""
    rng = np.random.default_rng(6)
    n, nb = 16, 7
    a, c = rng.uniform(size=(nb, n - 1)), rng.uniform(size=(nb, n - 1))
    b = 3 + rng.uniform(size=(nb, n))
    f = rng.uniform(size=(nb, n))
    x = O.thomas(a, b, c, f)
    for s in range(nb):
        ab = np.zeros((3, n))
        ab[0, 1:], ab[1], ab[2, :-1] = c[s], b[s], a[s]
        np.testing.assert_allclose(x[s], scipy.linalg.solve_banded((1, 1), ab, f[s]), rtol=1e-12)


# ---------------------------------------------------------------------------------------------
# closures: closed forms for linear velocity fields (halos set to the same linear field)
# ---------------------------------------------------------------------------------------------
def _linear_state(orc, grid, G, g=None):
    """u_i = Σ_j G[i, j] x_j on every parent point (halos included), tracer c = g · x."""
    H = grid.H
    lo = [1 - H[0], 1 - H[1], 1 - H[2]]
    hi = [grid.Nx + H[0], grid.Ny + H[1], grid.Nz + H[2]]
    dx, dy, dz = (e / n for e, n in zip(grid.extent, grid.N))
    i = np.arange(lo[0], hi[0] + 1)
    j = np.arange(lo[1], hi[1] + 1)
    k = np.arange(lo[2], hi[2] + 2)
    xc, xf = (i - 0.5) * dx, (i - 1) * dx
    yc, yf = (j - 0.5) * dy, (j - 1) * dy
    zc, zf = (k - 0.5) * dz, (k - 1) * dz
    locs = {0: (xf, yc, zc), 1: (xc, yf, zc), 2: (xc, yc, zf)}
    for q in range(3):
        X, Y, Z = np.meshgrid(*locs[q], indexing="ij")
        orc.set_range(q, lo, [hi[0], hi[1], hi[2] + 1], G[q, 0] * X + G[q, 1] * Y + G[q, 2] * Z)
    if g is not None:
        X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
        orc.set_range(3, lo, [hi[0], hi[1], hi[2] + 1], g[0] * X + g[1] * Y + g[2] * Z)


def test_amd_closed_form_for_uniform_velocity_gradient():
    """AMD on an isotropic grid with a uniform gradient G (G_ia = ∂ₐuᵢ): ν = max(0, −C·4Δ²·tr(GᵀSG)/|G|²),
    κ = max(0, −C·4Δ²·gᵀGg/|g|²) (anisotropic_minimum_dissipation.jl:161-204)."""
    grid = om.RectilinearGrid(size=(6, 6, 6), extent=(6, 6, 6), topology=(om.Periodic, om.Periodic, om.Periodic))
    kw = dict(advection=om.WENO(), closure=om.AnisotropicMinimumDissipation(), tracers=("c",))
    cfg, keep = om.build_config(grid, **kw)
    orc = O.OracleModel(cfg, grid)
    rng = np.random.default_rng(7)
    for trial in range(5):
        G = rng.normal(size=(3, 3)) * 1e-2
        G[2, 2] = -(G[0, 0] + G[1, 1])
        g = rng.normal(size=3)
        _linear_state(orc, grid, G, g)
        orc.compute_closure()
        nu, kap = orc.get(om._abi.OM_FIELD_NU), orc.get(om._abi.OM_FIELD_KAPPA0)
        S = 0.5 * (G + G.T)
        r = np.trace(G.T @ S @ G)
        q = np.sum(G ** 2)
        C, D2 = 1 / 3, 4 * 1.0 ** 2
        nu_exact = max(0.0, -C * D2 * r / q)
        kap_exact = max(0.0, -C * D2 * (g @ G @ g) / (g @ g))
        assert np.allclose(nu, nu_exact, rtol=1e-10, atol=1e-16)
        assert np.allclose(kap, kap_exact, rtol=1e-10, atol=1e-16)


def test_amd_zero_strain_short_circuit():
    grid = om.RectilinearGrid(size=(4, 4, 4), extent=(4, 4, 4))
    orc = oracle_for(grid, dict(advection=om.WENO(), closure=om.AnisotropicMinimumDissipation(), tracers=("c",)), {})
    orc.compute_closure()
    assert np.all(orc.get(om._abi.OM_FIELD_NU) == 0)
    assert np.all(orc.get(om._abi.OM_FIELD_KAPPA0) == 0)


def test_smagorinsky_lilly_closed_form():
    """νₑ = ς C² Δ² √(2 ΣᵢⱼΣᵢⱼ), ς = √(1 − min(1, Cb N²⁺/Σ²)) (smagorinsky.jl:90-104, lilly_coefficient.jl:129-142)."""
    grid = om.RectilinearGrid(size=(6, 6, 6), extent=(6, 6, 6), topology=(om.Periodic, om.Periodic, om.Periodic))
    kw = dict(advection=om.WENO(), closure=om.SmagorinskyLilly(C=0.16, Cb=1.0), buoyancy=om.BuoyancyTracer(), tracers=("b",))
    cfg, keep = om.build_config(grid, **kw)
    orc = O.OracleModel(cfg, grid)
    rng = np.random.default_rng(8)
    for N2 in (0.0, 1e-5, 5e-5):
        G = rng.normal(size=(3, 3)) * 1e-2
        _linear_state(orc, grid, G, np.array([0.0, 0.0, N2]))
        orc.compute_closure()
        S = 0.5 * (G + G.T)
        S2 = np.sum(S ** 2)
        sig = np.sqrt(1 - min(1.0, max(N2, 0) / S2))
        nu_exact = sig * 0.16 ** 2 * 1.0 * np.sqrt(2 * S2)
        np.testing.assert_allclose(orc.get(om._abi.OM_FIELD_NU), nu_exact, rtol=1e-10)


# ---------------------------------------------------------------------------------------------
# halos
# ---------------------------------------------------------------------------------------------
def test_halo_fill_brute_force():
    """Periodic copies (whole parent extent, corners included), one-layer Flux copies, Value and
    Gradient linear extrapolation, impenetrable w faces (fill_halo_regions_*.jl; test_halo_regions.jl:22-41)."""
    grid, kw, init = wind_convection(N=(6, 5, 4))
    cfg, keep = om.build_config(grid, **kw)
    orc = O.OracleModel(cfg, grid)
    rng = np.random.default_rng(9)
    Nx, Ny, Nz = grid.N
    H = grid.H
    for q in range(5):
        orc.set(q, rng.normal(size=(Nx, Ny, Nz + (1 if q == 2 else 0))))
    orc.fill_halos(fill_normal=True)
    lo = [1 - H[0], 1 - H[1], 1 - H[2]]
    hi = [Nx + H[0], Ny + H[1], Nz + H[2]]
    for q in range(5):
        P = orc.get_range(q, lo, hi)
        I = P[H[0]:H[0] + Nx, H[1]:H[1] + Ny, :]
        # periodic in x and y over every k
        np.testing.assert_array_equal(P[:H[0]], P[Nx:Nx + H[0]])
        np.testing.assert_array_equal(P[Nx + H[0]:], P[H[0]:2 * H[0]])
        np.testing.assert_array_equal(P[:, :H[1]], P[:, Ny:Ny + H[1]])
        kz0, kzN = H[2] - 1, H[2] + Nz       # parent indices of k = 0 and k = Nz + 1
        if q == 2:
            assert np.all(I[:, :, H[2]] == 0) and np.all(I[:, :, H[2] + Nz] == 0)   # w[1] = w[Nz+1] = 0
            continue
        c1, cN = I[:, :, H[2]], I[:, :, H[2] + Nz - 1]
        bot, top = I[:, :, kz0], I[:, :, kzN]
        if q == 3:   # T: top Flux (copy), bottom Gradient 0.01 with Δzᶠ = 1
            np.testing.assert_array_equal(top, cN)
            np.testing.assert_allclose(bot, c1 - 0.01 * 1.0, rtol=0, atol=1e-15)
        else:
            np.testing.assert_array_equal(top, cN)
            np.testing.assert_array_equal(bot, c1)


# ---------------------------------------------------------------------------------------------
# whole-model invariants and analytic solutions
# ---------------------------------------------------------------------------------------------
def test_resting_stratified_fluid_stays_at_rest():
    """b = N²z, U = 0 with AMD and the hydrostatic anomaly: the state is an exact steady solution
    (test_dynamics.jl:261-352)."""
    grid = om.RectilinearGrid(size=(8, 8, 8), extent=(8, 8, 8))
    kw = dict(advection=om.WENO(), closure=om.AnisotropicMinimumDissipation(), buoyancy=om.BuoyancyTracer(), tracers=("b",))
    orc = oracle_for(grid, kw, {"b": lambda x, y, z: 1e-5 * z})
    b0 = orc.get(3)
    orc.time_step(10.0, 5)
    for q in range(3):
        assert np.max(np.abs(orc.get(q))) == 0
    assert np.max(np.abs(orc.get(3) - b0)) == 0


def test_taylor_green_vortex_oracle():
    """Taylor–Green vortex, ν = 1, N = 64, 10 steps: max relative error < 5e-6 (test_dynamics.jl:214-259)."""
    N = 64
    grid, kw, init = taylor_green(N=N)
    orc = oracle_for(grid, kw, init)
    dt = (1 / (10 * np.pi)) * (1.0 / N) ** 2
    orc.time_step(dt, 10)
    t = orc.clock()[0]
    yC, xC = grid.ynodes("c"), grid.xnodes("c")
    ue = (-np.sin(2 * np.pi * yC) * np.exp(-4 * np.pi ** 2 * t))[None, :, None]
    ve = (np.sin(2 * np.pi * xC) * np.exp(-4 * np.pi ** 2 * t))[:, None, None]
    assert np.max(np.abs((orc.get(0) - ue) / ue)) < 5e-6
    assert np.max(np.abs((orc.get(1) - ve) / ve)) < 5e-6


def test_incompressible_in_time():
    """max|∇·U| stays at rounding level after steps (test_time_stepping.jl:138-171, atol 5e-8)."""
    grid, kw, init = convection(N=(8, 8, 8), noise_velocity=1e-2)
    orc = oracle_for(grid, kw, init)
    orc.time_step(1.0, 3)
    assert np.max(np.abs(orc.divergence())) < 5e-8
    assert np.max(np.abs(orc.divergence())) < 1e-15


def test_tracer_conservation_with_noflux_walls():
    """Volume integral of a tracer with NoFlux walls and no sources is conserved."""
    grid, kw, init = wind_convection(N=(8, 8, 8))
    kw = dict(kw, boundary_conditions={})
    init = dict(init, S=35 + np.random.default_rng(10).uniform(-1e-2, 1e-2, (8, 8, 8)))
    orc = oracle_for(grid, kw, init)
    s0 = orc.get(4).sum()
    orc.time_step(1.0, 3)
    assert abs(orc.get(4).sum() - s0) <= 1e-13 * abs(s0)


# ---------------------------------------------------------------------------------------------
# advection orientation and order, Coriolis, equation of state
# ---------------------------------------------------------------------------------------------
def _tendency_of(grid, kw, fields, target):
    """Tendency Gⁿ of prognostic `target` for the given interior fields (halos filled by the oracle)."""
    cfg, keep = om.build_config(grid, **kw)
    orc = O.OracleModel(cfg, grid)
    names = prognostic_names(kw)
    for name, arr in fields.items():
        orc.set(names.index(name), arr)
    orc.fill_halos(fill_normal=False)
    orc.compute_tendencies()
    G = orc.get(om._abi.OM_FIELD_GN + names.index(target))
    orc.close()
    return G


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno_tracer_flux_is_upwind_at_a_step(axis):
    """A tracer step advected by a uniform positive velocity: the Z-WENO5 face value at the jump is
    the upwind value (weno_interpolants.jl:348-354; upwind_biased_advective_fluxes.jl:99-121), so the
    last cell before the jump keeps G = 0 and the first cell after it gains u/Δ.  A downwind bias or a
    shifted stencil fails this."""
    N = 16
    grid = om.RectilinearGrid(size=(N, N, N), extent=(N, N, N), topology=(om.Periodic,) * 3)
    kw = dict(advection=om.WENO(), tracers=("c",))
    shape = (N, N, N)
    idx = np.indices(shape)[axis]
    c = np.where(idx < N // 2, 1.0, 0.0)
    c[idx >= N - 2] = 1.0                         # second jump back to 1 at the wrap-around, far from the first
    vel = {("u", "v", "w")[axis]: np.ones((N, N, N))}
    G = _tendency_of(grid, kw, dict(vel, c=c), "c")
    last_one = np.take(G, N // 2 - 1, axis=axis)
    first_zero = np.take(G, N // 2, axis=axis)
    assert np.max(np.abs(last_one)) < 1e-6
    np.testing.assert_allclose(first_zero, 1.0, rtol=1e-6)


def test_weno_tracer_advection_fifth_order():
    """Uniform u advecting cell averages of sin(2πx/L): the flux-divergence error falls ~32× per grid
    doubling (fifth order of WENO5 reconstruction on smooth data)."""
    errs = []
    for N in (32, 64):
        L = 1.0
        grid = om.RectilinearGrid(size=(N, 1, 1), extent=(L, 1, 1), topology=(om.Periodic,) * 3, halo=(3, 3, 3))
        dx = L / N
        xf = np.arange(N + 1) * dx
        cav = (-np.cos(2 * np.pi * xf[1:]) + np.cos(2 * np.pi * xf[:-1])) / (2 * np.pi * dx)
        exact = -(np.sin(2 * np.pi * xf[1:]) - np.sin(2 * np.pi * xf[:-1])) / dx
        G = _tendency_of(grid, dict(advection=om.WENO(), tracers=("c",)),
                         {"u": np.ones((N, 1, 1)), "c": cav.reshape(N, 1, 1)}, "c")
        errs.append(np.max(np.abs(G[:, 0, 0] - exact)))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 4.5, (errs, rate)


def test_weno_momentum_cross_fluxes_are_upwind():
    """u jumps in y and is advected by uniform v > 0: the momentum flux ṽ·u (ṽ the Centered(4)
    interpolation of A·v, upwind_biased_advective_fluxes.jl:39-48) takes the upwind u at the jump."""
    N = 16
    grid = om.RectilinearGrid(size=(N, N, N), extent=(N, N, N), topology=(om.Periodic,) * 3)
    j = np.indices((N, N, N))[1]
    u = np.where(j < N // 2, 1.0, 0.0)
    u[j >= N - 2] = 1.0
    G = _tendency_of(grid, dict(advection=om.WENO()), {"u": u, "v": np.ones((N, N, N))}, "u")
    assert np.max(np.abs(G[:, N // 2 - 1, :])) < 1e-6
    np.testing.assert_allclose(G[:, N // 2, :], 1.0, rtol=1e-6)


def test_fplane_coriolis_sign_and_magnitude():
    """Uniform (U, V) on an FPlane: Gu = +f V, Gv = −f U (f_plane.jl:8-54, coriolis_schemes.jl:67-68;
    uniform fields carry no advective flux divergence)."""
    N = 6
    grid = om.RectilinearGrid(size=(N, N, N), extent=(N, N, N), topology=(om.Periodic,) * 3)
    f, U, V = 1e-4, 0.3, -0.7
    kw = dict(advection=om.WENO(), coriolis=om.FPlane(f=f))
    fields = {"u": np.full((N, N, N), U), "v": np.full((N, N, N), V)}
    np.testing.assert_allclose(_tendency_of(grid, kw, fields, "u"), f * V, rtol=1e-14)
    np.testing.assert_allclose(_tendency_of(grid, kw, fields, "v"), -f * U, rtol=1e-14)


def test_linear_eos_resting_state_and_buoyancy_sign():
    """T = T₀ + Γz, S = S₀ at rest with the hydrostatic anomaly is steady; without the anomaly the w
    tendency is the buoyancy b = g(αT − βS) at w points (linear_equation_of_state.jl:39-40,72-80)."""
    grid = om.RectilinearGrid(size=(6, 6, 8), extent=(6, 6, 8))
    kw = dict(advection=om.WENO(), buoyancy=om.SeawaterBuoyancy(equation_of_state=om.LinearEquationOfState(2e-4, 8e-4)),
              tracers=("T", "S"))
    init = {"T": lambda x, y, z: 20 + 0.01 * z, "S": lambda x, y, z: 35.0 + 0 * z}
    orc = oracle_for(grid, kw, init)
    orc.time_step(10.0, 3)
    for q in range(3):
        assert np.max(np.abs(orc.get(q))) < 1e-15
    # periodic z: no anomaly, so Gw = ℑz b
    gridp = om.RectilinearGrid(size=(4, 4, 4), extent=(4, 4, 4), topology=(om.Periodic,) * 3)
    T, S = 21.0, 34.0
    Gw = _tendency_of(gridp, kw, {"T": np.full((4, 4, 4), T), "S": np.full((4, 4, 4), S)}, "w")
    b = 9.80665 * (2e-4 * T - 8e-4 * S)
    np.testing.assert_allclose(Gw[:, :, :4], b, rtol=1e-13)


# ---------------------------------------------------------------------------------------------
# conjugate-gradient Poisson solver
# ---------------------------------------------------------------------------------------------
def _volume_laplacian_matrix(grid, dzc, dzf):
    """Sparse V∇² of the cell-centred pressure: periodic x, y; Neumann (no coupling) across Bounded
    z walls; face coefficients A/Δ (conjugate_gradient_poisson_solver.jl:62-66)."""
    import scipy.sparse as sp
    Nx, Ny, Nz = grid.N
    dx, dy = grid.extent[0] / Nx, grid.extent[1] / Ny
    idx = lambda i, j, k: (i % Nx) + Nx * ((j % Ny) + Ny * k)
    rows, cols, vals = [], [], []
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                c = idx(i, j, k)
                ax, ay = dy * dzc[k] / dx, dx * dzc[k] / dy
                for ii, jj, a in [(i - 1, j, ax), (i + 1, j, ax), (i, j - 1, ay), (i, j + 1, ay)]:
                    rows += [c, c]; cols += [idx(ii, jj, k), c]; vals += [a, -a]
                for kk, face in [(k - 1, k), (k + 1, k + 1)]:
                    if grid.topology[2] is om.Bounded and not 0 <= kk < Nz:
                        continue
                    a = dx * dy / dzf[face]
                    rows += [c, c]; cols += [idx(i, j, kk % Nz), c]; vals += [a, -a]
    return sp.csr_matrix((vals, (rows, cols)), shape=(Nx * Ny * Nz,) * 2)


CG_GRIDS = {
    "PPB": dict(z=(-8, 0), topology=(om.Periodic, om.Periodic, om.Bounded)),
    "PPP": dict(z=(-8, 0), topology=(om.Periodic, om.Periodic, om.Periodic)),
    "PPB_stretched": dict(z=[-8, -6.5, -5, -3.8, -2.5, -1.5, -0.8, -0.3, 0], topology=(om.Periodic, om.Periodic, om.Bounded)),
}


@pytest.mark.parametrize("gname", sorted(CG_GRIDS))
@pytest.mark.parametrize("precond", ["default", "none", "fft", "ct"])
def test_conjugate_gradient_matches_sparse_direct_solve(gname, precond):
    """The CG solution of V∇²x = V·rhs (zero-mean gauge) equals scipy's sparse direct solve of the
    same operator; the FFT / Fourier–tridiagonal preconditioner is the exact inverse (one iteration)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    grid = om.RectilinearGrid(size=(8, 6, 8), x=(0, 8), y=(0, 6), **CG_GRIDS[gname])
    pc = {"default": "default", "none": None,
          "fft": om.FourierTridiagonalPoissonSolver() if grid.stretched else om.FFTBasedPoissonSolver(),
          "ct": om.ColumnwiseTridiagonalPreconditioner()}[precond]
    cfg, keep = om.build_config(grid, advection=om.WENO(),
                                pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=pc, reltol=1e-14, abstol=0))
    o = O.OracleModel(cfg, grid)
    dzc, dzf, zc, zf = o.grid_z()
    H = grid.H[2]
    dzc_i = dzc[H:H + grid.Nz]
    dzf_i = dzf[H:H + grid.Nz + 1]
    A = _volume_laplacian_matrix(grid, dzc_i, dzf_i)
    V = (grid.extent[0] / 8) * (grid.extent[1] / 6) * np.repeat(dzc_i, 8 * 6)
    rf = np.random.default_rng(11).normal(size=8 * 6 * 8)
    rf -= (V * rf).sum() / V.sum()                       # compatibility with the Neumann / periodic null space
    n = A.shape[0]
    aug = sp.bmat([[A, np.ones((n, 1))], [np.ones((1, n)), None]]).tocsc()
    exact = spl.spsolve(aug, np.concatenate([V * rf, [0.0]]))[:n].reshape((8, 6, 8), order="F")
    phi = o.poisson_solve(rf.reshape((8, 6, 8), order="F"))
    it, res, true_res = o.last_solve_info()
    assert abs(phi.mean()) < 1e-13 * np.abs(exact).max()
    assert np.abs(phi - exact).max() <= 1e-12 * np.abs(exact).max()
    assert true_res <= 1e-12 * np.linalg.norm(rf)
    if precond == "fft":
        assert it == 1


def test_columnwise_tridiagonal_preconditioner_inverts_the_vertical_operator():
    """One application of ColumnwiseTridiagonalPreconditioner solves, in every column, the system of
    the reference's docstring (conjugate_gradient_poisson_solver.jl:349-363):
    Az⁻ p[k−1] − (Az⁻ + Az⁺)(1 + 1/100) p[k] + Az⁺ p[k+1] = r[k], Az± = Az/Δzᶠ at the cell's lower and
    upper faces, no coupling through the Bounded walls or into immersed cells, identity rows for
    immersed and vertically isolated cells; checked against scipy's banded solver on the stretched
    seamount grid."""
    from scipy.linalg import solve_banded
    grid, kw, _ = _seamount_grid()
    cfg, keep = om.build_config(grid, advection=om.WENO(),
                                pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=om.ColumnwiseTridiagonalPreconditioner()))
    o = O.OracleModel(cfg, grid)
    imm = _immersed_cells(grid, o)
    Nx, Ny, Nz = grid.N
    dzc, dzf, zc, zf = o.grid_z()
    dzf_i = dzf[grid.H[2]:grid.H[2] + Nz + 1]
    Az = (grid.extent[0] / Nx) * (grid.extent[1] / Ny)
    r = np.random.default_rng(5).normal(size=grid.N)
    z = o.cg_precondition(om._abi.OM_PRECOND_COLUMNWISE_TRIDIAGONAL, r)
    for i in range(Nx):
        for j in range(Ny):
            act = ~imm[i, j, :]
            ab = np.zeros((3, Nz))
            for k in range(Nz):
                if not act[k]:
                    ab[1, k] = 1.0
                    continue
                below = k > 0 and act[k - 1]
                above = k < Nz - 1 and act[k + 1]
                am = Az / dzf_i[k] if below else 0.0
                ap = Az / dzf_i[k + 1] if above else 0.0
                ab[1, k] = 1.0 if not (below or above) else -(am + ap) * 1.01
                if above:
                    ab[0, k + 1] = ap
                    ab[2, k] = ap
            x = solve_banded((1, 1), ab, r[i, j, :])
            assert np.abs(z[i, j, :] - x).max() <= 1e-13 * np.abs(x).max()


def test_columnwise_tridiagonal_preconditioner_on_a_flat_aspect_grid():
    """(Δz/Δx)²Nz² ≪ 1: the columnwise preconditioner needs several times fewer CG iterations than
    none, as the reference's docstring states (conjugate_gradient_poisson_solver.jl:365-368), and
    reaches the same solution."""
    grid = om.RectilinearGrid(size=(16, 16, 8), x=(0, 1), y=(0, 1), z=(-1 / 64, 0))
    rhs = np.random.default_rng(9).normal(size=grid.N)
    rhs -= rhs.mean()
    out = {}
    for name, pc in (("none", None), ("ct", om.ColumnwiseTridiagonalPreconditioner())):
        cfg, keep = om.build_config(grid, advection=om.WENO(),
                                    pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=pc, reltol=1e-12, abstol=0))
        o = O.OracleModel(cfg, grid)
        out[name] = (o.poisson_solve(rhs), o.last_solve_info()[0])
    print("iterations none / columnwise", out["none"][1], out["ct"][1])
    assert 3 * out["ct"][1] <= out["none"][1]
    assert np.abs(out["ct"][0] - out["none"][0]).max() <= 1e-9 * np.abs(out["none"][0]).max()


# ---------------------------------------------------------------------------------------------
# immersed bottom (GridFittedBottom)
# ---------------------------------------------------------------------------------------------
def _seamount_grid(N=(12, 10, 8), stretched=True):
    from cases import seamount
    grid, kw, init = seamount(N=N, stretched=stretched)
    return grid, kw, init


def test_immersed_resting_stratified_fluid_stays_at_rest():
    """b = N²z at rest over a grid-fitted seamount with the hydrostatic anomaly: the conditional
    differences leave no spurious pressure-gradient force, so the state is exactly steady
    (the immersed counterpart of test_dynamics.jl:261-352)."""
    grid, kw, _ = _seamount_grid()
    orc = oracle_for(grid, kw, {"b": lambda x, y, z: 1e-6 * z})
    b0 = orc.get(3)
    orc.time_step(30.0, 3)
    for q in range(3):
        assert np.max(np.abs(orc.get(q))) == 0
    np.testing.assert_array_equal(orc.get(3), b0)


def _immersed_cells(grid, o):
    """Boolean (Nx, Ny, Nz) immersed mask from the numerical bottom rule (zᶜ ≤ z_b)."""
    dzc, dzf, zc, zf = o.grid_z()
    H = grid.H[2]
    zcent = zc[H:H + grid.Nz]
    zb = grid.bottom_heights()
    return zcent[None, None, :] <= zb[:, :, None]


def test_immersed_cg_matches_sparse_direct_solve():
    """CG on V∇² with the conditional (Neumann) faces next to immersed cells equals scipy's direct
    solve of the same operator restricted to the active cells; immersed cells hold 0 (the gauge)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    grid, kw, _ = _seamount_grid()
    cfg, keep = om.build_config(grid, advection=om.WENO(),
                                pressure_solver=om.ConjugateGradientPoissonSolver(reltol=1e-14, abstol=0))
    o = O.OracleModel(cfg, grid)
    imm = _immersed_cells(grid, o)
    assert imm.any() and not imm.all()
    Nx, Ny, Nz = grid.N
    dzc, dzf, zc, zf = o.grid_z()
    H = grid.H[2]
    dzc_i, dzf_i = dzc[H:H + Nz], dzf[H:H + Nz + 1]
    dx, dy = grid.extent[0] / Nx, grid.extent[1] / Ny
    act = ~imm
    ids = -np.ones((Nx, Ny, Nz), dtype=int)
    ids[act] = np.arange(act.sum())
    rows, cols, vals = [], [], []
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                if not act[i, j, k]:
                    continue
                c = ids[i, j, k]
                nb = [((i - 1) % Nx, j, k, dy * dzc_i[k] / dx), ((i + 1) % Nx, j, k, dy * dzc_i[k] / dx),
                      (i, (j - 1) % Ny, k, dx * dzc_i[k] / dy), (i, (j + 1) % Ny, k, dx * dzc_i[k] / dy)]
                if k > 0:
                    nb.append((i, j, k - 1, dx * dy / dzf_i[k]))
                if k < Nz - 1:
                    nb.append((i, j, k + 1, dx * dy / dzf_i[k + 1]))
                for a, b, cc, coef in nb:
                    if act[a, b, cc]:
                        rows += [c, c]; cols += [ids[a, b, cc], c]; vals += [coef, -coef]
    n = int(act.sum())
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    V = dx * dy * np.broadcast_to(dzc_i[None, None, :], (Nx, Ny, Nz))
    rhs = np.random.default_rng(12).normal(size=(Nx, Ny, Nz))
    rhs[imm] = 0
    rhs[act] -= (V[act] * rhs[act]).sum() / V[act].sum()
    aug = sp.bmat([[A, np.ones((n, 1))], [np.ones((1, n)), None]]).tocsc()
    x = spl.spsolve(aug, np.concatenate([V[act] * rhs[act], [0.0]]))[:n]
    phi = o.poisson_solve(rhs)
    it, res, true_res = o.last_solve_info()
    assert np.all(phi[imm] == 0)
    assert np.abs(phi[act] - phi[act].mean() - (x - x.mean())).max() <= 1e-11 * np.abs(x).max()
    assert it >= 2   # the FFT preconditioner of the underlying grid is no longer exact


def test_immersed_projection_and_tracer_conservation():
    """After steps over the seamount: velocities vanish on the immersed periphery, the active cells
    are divergence-free, and Σ V b over the fluid is conserved (no-flux immersed and wall faces)."""
    grid, kw, init = _seamount_grid()
    init = dict(init, b=lambda x, y, z: 1e-6 * z + 1e-8 * np.sin(x / 700.0) * np.cos(y / 900.0))
    orc = oracle_for(grid, kw, init)
    cfg_o = O.OracleModel(*(om.build_config(grid, **kw)[0], grid))
    imm = _immersed_cells(grid, cfg_o)
    dzc = cfg_o.grid_z()[0][grid.H[2]:grid.H[2] + grid.Nz]
    V = np.broadcast_to(dzc[None, None, :], grid.N)
    total0 = (V * orc.get(3))[~imm].sum()
    orc.time_step(20.0, 3)
    u, v, w = orc.get(0), orc.get(1), orc.get(2)
    # faces between an immersed cell and its neighbour (x): u at face i touches cells i-1 and i
    imm_u = imm | np.roll(imm, 1, axis=0)
    imm_v = imm | np.roll(imm, 1, axis=1)
    assert np.all(u[imm_u] == 0) and np.all(v[imm_v] == 0)
    imm_w = np.zeros(w.shape, dtype=bool)
    imm_w[:, :, 1:grid.Nz] = imm[:, :, 1:] | imm[:, :, :-1]
    assert np.all(w[imm_w] == 0)
    div = orc.divergence()
    assert np.abs(div[~imm]).max() <= 1e-10 * np.abs(u).max() / (grid.extent[0] / grid.Nx)
    total = (V * orc.get(3))[~imm].sum()
    assert abs(total - total0) <= 1e-12 * abs(total0)


# ---------------------------------------------------------------------------------------------
# cell_advection_timescale (src/Advection/cell_advection_timescale.jl:14-35)
# ---------------------------------------------------------------------------------------------
def test_cell_advection_timescale_closed_forms():
    """Uniform (U, V, W): τ = 1/(|U|/Δx + |V|/Δy + |W|/Δz) (w vanishes on the bottom wall, which
    only removes one layer of cells from the minimum); on a stretched grid the thinnest face
    spacing sets τ; a fluid at rest gives +Inf; Flat directions contribute nothing."""
    grid = om.RectilinearGrid(size=(4, 5, 6), x=(0, 2), y=(0, 5), z=(-3, 0))
    U, V, W = 0.3, -0.2, 0.05
    orc = oracle_for(grid, {}, {})
    for q, val in enumerate((U, V, W)):
        orc.set(q, np.full(om.models.interior_shape(grid, q), val))
    expect = 1.0 / (abs(U) / 0.5 + abs(V) / 1.0 + abs(W) / 0.5)
    assert orc.cell_advection_timescale() == pytest.approx(expect, rel=1e-15)
    rest = oracle_for(grid, {}, {})
    assert rest.cell_advection_timescale() == np.inf
    sgrid = om.RectilinearGrid(size=(4, 4, 4), x=(0, 1), y=(0, 1), z=[-4, -2, -1, -0.25, 0])
    so = oracle_for(sgrid, {}, {})
    so.set(2, np.full(om.models.interior_shape(sgrid, 2), 1.0))
    dzf = so.grid_z()[1][sgrid.H[2]:sgrid.H[2] + 5]            # face spacings, faces 1..5
    assert so.cell_advection_timescale() == pytest.approx(dzf[0:4].min(), rel=1e-15)
    g2 = om.RectilinearGrid(size=(8, 8), x=(0, 1), y=(0, 2), topology=(om.Periodic, om.Periodic, om.Flat))
    o2 = oracle_for(g2, {"advection": om.WENO()}, {"u": 2.0, "v": 1.0})
    assert o2.cell_advection_timescale() == pytest.approx(1.0 / (2.0 / 0.125 + 1.0 / 0.25), rel=1e-15)


def test_fp32_oracle_field_round_trip():
    """An FP32 oracle stores set! values rounded to FP32 and returns them exactly."""
    grid = om.RectilinearGrid(float_type=np.float32, size=(4, 3, 5), extent=(1, 1, 1))
    cfg, keep = om.build_config(grid)
    o = O.OracleModel(cfg, grid)
    a = np.random.default_rng(4).normal(size=(4, 3, 5))
    o.set(3 - 3, a)
    np.testing.assert_array_equal(o.get(0), a.astype(np.float32).astype(np.float64))

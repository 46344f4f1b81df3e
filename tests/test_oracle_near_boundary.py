"""Pins of the oracle where the reference tree holds no vectors: the composite near a Bounded z
wall, pHY′'s column integral, the AMD interpolation placements, the immersed reduced-order
stencils and Value boundary conditions.  Where a pin could pass on a plausible mistake, the
same check is run on an oracle build with that mistake planted (om_oracle.cpp ORC_PLANT,
oracle/liboracle_plant<N>.so) and must fail there.

Reference anchors:
  src/Advection/topologically_conditional_interpolation.jl:44-58,106-124 (buffer bands)
  src/Advection/weno_reconstruction.jl:107-140 (WENO5 → WENO3 → Centered(2) chain)
  src/Models/NonhydrostaticModels/update_hydrostatic_pressure.jl:11-21
  src/TurbulenceClosures/turbulence_closure_implementations/anisotropic_minimum_dissipation.jl:161-358
  test/test_immersed_advection.jl:22-90 (constant masked tracer reconstructs to 1)
  src/BoundaryConditions/fill_halo_regions_value_gradient.jl:7-60
"""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import convection, oracle_for, prognostic_names, stretched_convection, wind_convection
from oracle import oracle as O

A = om._abi


# ---------------------------------------------------------------------------------------------
# near a Bounded z wall
# ---------------------------------------------------------------------------------------------
def _tendencies(orc, nf):
    return [orc.get(A.OM_FIELD_GN + q).copy() for q in range(nf)]


def _set_z_halo_layer(orc, grid, fid, layer, value, face=False):
    """Set z-halo layer `layer` (1 = next to the wall) on both sides of a field, whole x-y extent."""
    Nx, Ny, Nz = grid.N
    H = grid.H
    for k in (1 - layer, Nz + (1 if face else 0) + layer):
        orc.set_range(fid, (1 - H[0], 1 - H[1], k), (Nx + H[0], Ny + H[1], k), value)


@pytest.mark.parametrize("case", ["convection", "wind_convection", "stretched_convection"])
def test_near_wall_stencils_reach_only_the_first_halo_layer(case):
    """With WENO5 → WENO3 → Centered(2) switching at the buffer bands, every z stencil of a tendency
    stays within one halo layer of the wall: NaN in z-halo layers 2 and 3 of every prognostic and
    closure field, and in all three layers of pHY′ (read only by horizontal differences), leaves the
    tendencies bit for bit unchanged.  (Layer 1 is read: one-layer Flux/Value/Gradient halos.)"""
    builders = {"convection": lambda: convection(N=(12, 10, 14), noise_velocity=1e-2),
                "wind_convection": lambda: wind_convection(N=(12, 10, 14)),
                "stretched_convection": lambda: stretched_convection(N=(12, 10, 14))}
    grid, kw, init = builders[case]()
    names = prognostic_names(kw)
    nf, nt = len(names), len(names) - 3
    orc = oracle_for(grid, kw, init)
    orc.compute_tendencies()
    G0 = _tendencies(orc, nf)
    assert all(np.isfinite(g).all() for g in G0)
    aux = [A.OM_FIELD_NU] if kw.get("closure") is not None else []
    if isinstance(kw.get("closure"), om.AnisotropicMinimumDissipation):
        aux += [A.OM_FIELD_KAPPA0 + n for n in range(nt)]
    for fid in list(range(nf)) + aux:
        for layer in (2, 3):
            _set_z_halo_layer(orc, grid, fid, layer, np.nan, face=(fid == 2))
    if kw.get("buoyancy") is not None:
        for layer in (1, 2, 3):
            _set_z_halo_layer(orc, grid, A.OM_FIELD_PHY, layer, np.nan)
    orc.compute_tendencies()
    for q, (a, b) in enumerate(zip(G0, _tendencies(orc, nf))):
        assert np.array_equal(a, b), f"{case}: tendency of {names[q]} read a z-halo layer beyond the first"
    # the first layer is read: NaN there reaches the tracer tendency
    _set_z_halo_layer(orc, grid, 3, 1, np.nan)
    orc.compute_tendencies()
    assert not np.isfinite(orc.get(A.OM_FIELD_GN + 3)).all()


def _band_violations(plant, W0, field):
    """Face values of a z reconstruction recovered from the advective tendency of a field that
    varies only in z, under uniform vertical flow W0 (w = 0 on the walls), checked against the
    band rule: WENO5 on faces 4..Nz-2, WENO3 on 3 and Nz-1, Centered(2) on 2 and Nz."""
    Nz = 16
    grid = om.RectilinearGrid(size=(4, 4, Nz), x=(0, 4), y=(0, 4), z=(-Nz, 0))
    cfg, keep = om.build_config(grid, advection=om.WENO(), tracers=("c",))
    m = O.OracleModel(cfg, grid, plant=plant)
    k = np.arange(1, Nz + 1, dtype=float)
    a, b, c0 = 0.37, -1.3, 2.0
    avg = a * (k * k - k + 1.0 / 3.0) + b * (k - 0.5) + c0      # cell averages of a z² + b z + c0, cell k on [k-1, k]
    q = 0 if field == "u" else 3
    prof = np.broadcast_to(avg[None, None, :], (4, 4, Nz))
    m.set(q, prof)
    m.set(2, np.full((4, 4, Nz + 1), W0))
    m.fill_halos(fill_normal=True)                               # w[1] = w[Nz+1] = 0
    m.compute_tendencies()
    G = m.get(A.OM_FIELD_GN + q)[0, 0, :]
    # G[k] = -(F[k+1] - F[k]) / Δz with F = W̃ · value at the face and F[1] = 0 (Δz = 1)
    Wt = W0   # the advecting velocity: w itself for c, the Centered(4) x-average of Az·w (Az = 1) for u
    face = {}
    flux = 0.0
    for kk in range(1, Nz):                                      # faces 2 .. Nz
        flux = flux - G[kk - 1]
        face[kk + 1] = flux / Wt
    left = W0 > 0
    bad = []
    for f in range(2, Nz + 1):
        if 4 <= f <= Nz - 2:
            want = O.weno5(avg[f - 4:f + 2], left=left)
            exact = a * (f - 1) ** 2 + b * (f - 1) + c0
            if abs(want - exact) > 1e-12 * abs(exact):               # WENO5 reproduces quadratics
                bad.append((f, "weno5 not exact"))
        elif f in (3, Nz - 1):
            want = O.weno3(avg[f - 3:f + 1], left=left)
        else:
            want = 0.5 * avg[f - 2] + 0.5 * avg[f - 1]
        if abs(face[f] - want) > 1e-11 * max(1.0, abs(want)):
            bad.append((f, face[f], want))
    return bad


@pytest.mark.parametrize("field", ["c", "u"])
@pytest.mark.parametrize("W0", [1.0, -1.0])
def test_near_wall_scheme_bands(field, W0):
    assert _band_violations(0, W0, field) == []


def test_near_wall_scheme_bands_fail_on_a_widened_band():
    """The same check on an oracle that uses WENO5 one face too close to the bottom wall fails."""
    assert _band_violations(3, 1.0, "c") != []


def _phy_mismatch(plant):
    """pHY′ for a buoyancy that varies only horizontally, b = f(x, y), on a stretched z:
    pHY′[k] = −f Σ_{k' = k+1}^{Nz+1} Δzᶠ(k'), summed from the top (the halo b[Nz+1] = b[Nz])."""
    grid, kw, _ = stretched_convection(N=(8, 6, 16))
    Lx, Ly = grid.extent[0], grid.extent[1]
    init = {"b": lambda x, y, z: 2e-3 + 1e-3 * (np.sin(2 * np.pi * x / Lx) + 0.5 * np.cos(2 * np.pi * y / Ly)) + 0 * z}
    cfg, keep = om.build_config(grid, **kw)
    o = O.OracleModel(cfg, grid, plant=plant)
    o.set(3, om.models._evaluate(grid, "b", init["b"]))
    o.set_finalize(True)
    b = o.get(3)
    f = b[:, :, -1]
    assert np.array_equal(b, np.broadcast_to(f[:, :, None], b.shape))
    dzc, dzf, zc, zf = o.grid_z()                     # k = 1-Hz .. Nz+Hz
    Hz, Nz = grid.H[2], grid.Nz
    dzF = lambda k: dzf[k - (1 - Hz)]
    want = np.empty(b.shape)
    p = -f * dzF(Nz + 1)
    want[:, :, Nz - 1] = p
    for k in range(Nz - 1, 0, -1):
        p = p - f * dzF(k + 1)
        want[:, :, k - 1] = p
    got = o.get(A.OM_FIELD_PHY)
    return np.abs(got - want).max() / np.abs(want).max()


def test_hydrostatic_pressure_closed_form():
    assert _phy_mismatch(0) <= 1e-15


def test_hydrostatic_pressure_closed_form_fails_on_the_wrong_top_spacing():
    assert _phy_mismatch(2) > 1e-6


# ---------------------------------------------------------------------------------------------
# AMD interpolation placements: second-order convergence to the continuum closure
# ---------------------------------------------------------------------------------------------
TP = 2 * np.pi


def _U(x, y, z): return np.sin(x + 0.3) * np.cos(2 * y) + 0.5 * np.cos(z - 0.2) * np.sin(y) + 0.2 * np.sin(x + z)
def _V(x, y, z): return np.cos(x) * np.sin(y + 0.1) * np.cos(z) + 0.3 * np.sin(2 * x + 0.5)
def _W_flat_y(x, y, z): return 0.7 * np.sin(z + 0.4) * np.cos(x - 0.3) + 0.4 * np.cos(2 * x + z)
def _W(x, y, z): return _W_flat_y(x, y, z) + 0.8 * np.sin(y + z) * np.cos(x)
def _C(x, y, z): return np.sin(x) + 0.6 * np.cos(y + 0.7) * np.sin(z) + 0.5 * np.cos(x + y - z)


def _grad(f, x, y, z, e=1e-3):
    """Fourth-order central differences of an analytic field (error ~1e-12)."""
    out = []
    for d in range(3):
        def sh(s):
            a = [x, y, z]
            a[d] = a[d] + s * e
            return f(*a)
        out.append((8 * (sh(1) - sh(-1)) - (sh(2) - sh(-2))) / (12 * e))
    return out


def _amd_errors(N, W, plant=0):
    """RMS errors of νₑ/(Cν δ²) and κₑ/(Cκ δ²) against the continuum AMD closure
    −(∂ₖuᵢ ∂ₖuⱼ Sᵢⱼ)/(∂ₗuₘ ∂ₗuₘ) and −(Sᵢⱼ ∂ᵢc ∂ⱼc)/|∇c|² (the filter widths cancel on a cubic
    grid), over the cells where the denominators are at least 0.3 of their maximum (where the
    ratio is well conditioned) and the closure is active."""
    grid = om.RectilinearGrid(size=(N, N, N), x=(0, TP), y=(0, TP), z=(0, TP),
                              topology=(om.Periodic, om.Periodic, om.Periodic))
    cfg, keep = om.build_config(grid, closure=om.AnisotropicMinimumDissipation(), tracers=("c",))
    m = O.OracleModel(cfg, grid, plant=plant)
    h = TP / N
    c = (np.arange(N) + 0.5) * h
    f = np.arange(N) * h
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    m.set(0, _U(*np.meshgrid(f, c, c, indexing="ij")))
    m.set(1, _V(*np.meshgrid(c, f, c, indexing="ij")))
    m.set(2, W(*np.meshgrid(c, c, f, indexing="ij")))
    m.set(3, _C(X, Y, Z))
    m.fill_halos()
    m.compute_closure()
    d2 = 3.0 / (3.0 / (2 * h) ** 2)
    Fn = m.get(A.OM_FIELD_NU) / (d2 / 3)
    Fk = m.get(A.OM_FIELD_KAPPA0) / (d2 / 3)
    G = [_grad(g, X, Y, Z) for g in (_U, _V, W)]          # G[i][k] = ∂ₖuᵢ
    S = [[0.5 * (G[i][j] + G[j][i]) for j in range(3)] for i in range(3)]
    r = sum(G[i][k] * G[j][k] * S[i][j] for i in range(3) for j in range(3) for k in range(3))
    q = sum(G[i][k] ** 2 for i in range(3) for k in range(3))
    gc = _grad(_C, X, Y, Z)
    th = sum(S[i][j] * gc[i] * gc[j] for i in range(3) for j in range(3))
    s = sum(g * g for g in gc)
    cn, ck = -r / q, -th / s
    seln = (q > 0.3 * q.max()) & (cn > 0.05)
    selk = (s > 0.3 * s.max()) & (ck > 0.05)
    rms = lambda e: float(np.sqrt(np.mean(e ** 2)))
    return rms((Fn - cn)[seln]), rms((Fk - ck)[selk])


def _orders(W, plant=0):
    e = [_amd_errors(N, W, plant) for N in (16, 32, 64)]
    nu = [np.log2(e[q][0] / e[q + 1][0]) for q in range(2)]
    kap = [np.log2(e[q][1] / e[q + 1][1]) for q in range(2)]
    print("AMD orders", W.__name__, "plant", plant, "nu", nu, "kappa", kap, flush=True)
    return nu, kap


def test_amd_converges_at_second_order_to_the_continuum_closure():
    """With w independent of y, every term of νₑ and κₑ is interpolated to the cell centre, and both
    converge at second order (16³ → 32³ → 64³)."""
    nu, kap = _orders(_W_flat_y)
    assert min(nu) >= 1.85 and min(kap) >= 1.85


def test_amd_diffusivity_keeps_the_reference_placement_of_dy_w():
    """The reference interpolates ∂y w (a cff quantity) to the centre with ℑxzᶜᵃᶜ
    (anisotropic_minimum_dissipation.jl:343), half a cell off in x and y: with ∂y w ≠ 0 the
    diffusivity converges at first order asymptotically, while the viscosity keeps second order."""
    nu, kap = _orders(_W)
    assert min(nu) >= 1.85
    assert kap[1] < 1.6 and kap[1] < kap[0]


def test_amd_convergence_fails_on_a_misplaced_interpolation():
    """One ffc term of νₑ averaged with ℑxz instead of ℑxy: the order drops below the bar."""
    nu, _ = _orders(_W_flat_y, plant=1)
    assert nu[1] < 1.85


# ---------------------------------------------------------------------------------------------
# immersed bottom: a constant masked tracer (test/test_immersed_advection.jl:22-90)
# ---------------------------------------------------------------------------------------------
def test_immersed_constant_tracer_reconstructs_to_one():
    """c = 1 over a grid-fitted seamount, masked to 0 in the immersed cells: every reduced-order
    reconstruction next to the bottom must use active cells only, so the advective tendency of c
    is −c ∇·U to rounding (∇·U = 0 after set!'s projection) and c stays 1 over steps."""
    from cases import seamount
    grid, kw, init = seamount(N=(12, 10, 8))
    kw = dict(kw, tracers=("b", "c"))
    init = dict(init, c=1.0)
    orc = oracle_for(grid, kw, init)
    cfg, keep = om.build_config(grid, **kw)
    dzc, dzf, zc, zf = O.OracleModel(cfg, grid).grid_z()
    H = grid.H[2]
    imm = zc[H:H + grid.Nz][None, None, :] <= grid.bottom_heights()[:, :, None]
    assert imm.any()
    c = orc.get(4)
    assert np.all(c[imm] == 0) and np.all(c[~imm] == 1)
    Gc = orc.get(A.OM_FIELD_GN + 4)
    scale = 0.1 / (grid.extent[0] / grid.Nx)                     # |U| / Δx
    assert np.abs(Gc[~imm]).max() <= 1e-9 * scale
    orc.time_step(20.0, 3)
    c = orc.get(4)
    assert np.abs(c[~imm] - 1).max() <= 1e-9


# ---------------------------------------------------------------------------------------------
# Value boundary condition: c_halo = 2 v − c₁
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("stretched", [False, True])
def test_value_boundary_condition_halo_closed_form(stretched):
    """A Value condition v on a Center field extrapolates linearly through the boundary face:
    with ∇c = (c₁ − v)/(Δ/2) and c₀ = c₁ − ∇c Δ the halo is 2v − c₁, whatever Δ is."""
    if stretched:
        grid = om.RectilinearGrid(size=(6, 5, 5), x=(0, 6), y=(0, 5), z=[-7.0, -4.5, -2.6, -1.3, -0.5, 0.0])
    else:
        grid = om.RectilinearGrid(size=(6, 5, 4), x=(0, 6), y=(0, 5), z=(-4, 0))
    vb, vt = 0.7, -1.9
    bcs = {"c": om.FieldBoundaryConditions(bottom=om.ValueBoundaryCondition(vb), top=om.ValueBoundaryCondition(vt)),
           "u": om.FieldBoundaryConditions(top=om.ValueBoundaryCondition(0.25))}
    cfg, keep = om.build_config(grid, tracers=("c",), boundary_conditions=bcs)
    o = O.OracleModel(cfg, grid)
    rng = np.random.default_rng(17)
    Nx, Ny, Nz = grid.N
    o.set(3, rng.normal(size=(Nx, Ny, Nz)))
    o.set(0, rng.normal(size=(Nx, Ny, Nz)))
    o.fill_halos(fill_normal=True)
    for q, bot, top in ((3, vb, vt), (0, None, 0.25)):
        c1 = o.get_range(q, (1, 1, 1), (Nx, Ny, 1))
        cN = o.get_range(q, (1, 1, Nz), (Nx, Ny, Nz))
        h0 = o.get_range(q, (1, 1, 0), (Nx, Ny, 0))
        hN = o.get_range(q, (1, 1, Nz + 1), (Nx, Ny, Nz + 1))
        if bot is not None:
            np.testing.assert_allclose(h0, 2 * bot - c1, rtol=0, atol=8 * np.finfo(float).eps * 4)
        else:
            np.testing.assert_array_equal(h0, c1)
        np.testing.assert_allclose(hN, 2 * top - cN, rtol=0, atol=8 * np.finfo(float).eps * 4)

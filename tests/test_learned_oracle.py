"""Pins of the learned-preconditioner oracle (oracle/learned.py): BF16 rounding, the convolution
and its adjoint, the gradients, and the Adam step.  The reference has no such network, so every
pin is a closed form, a brute-force evaluation or a finite difference."""
import numpy as np
import pytest

from oracle import learned as L


def test_bf16_rounding_matches_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(0).normal(size=10000).astype(np.float32) * np.float32(3.7)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(L.bf16(x), ref)
    one = np.float32(1.0)
    ties = np.array([one + np.float32(2.0 ** -8), one + np.float32(3 * 2.0 ** -8), one + np.float32(2.0 ** -9)], dtype=np.float32)
    np.testing.assert_array_equal(L.bf16(ties), np.array([1.0, 1.0 + 2.0 ** -6, 1.0], dtype=np.float32))


def test_convolution_against_brute_force():
    """y[i,j,k,o] = Σ x[(i+dx) mod Nx, (j+dy) mod Ny, k+dz, c] W[t,c,o], zero outside 0 ≤ k+dz < Nz."""
    rng = np.random.default_rng(1)
    Nx, Ny, Nz = 3, 4, 3
    x = rng.normal(size=(Nx, Ny, Nz, 16))
    w = rng.normal(size=(27, 16, 16))
    y = L.conv(x, w, np.float64)
    ref = np.zeros_like(y)
    for i in range(Nx):
        for j in range(Ny):
            for k in range(Nz):
                for t, (dz, dy, dx) in enumerate(L.taps()):
                    kk = k + dz
                    if 0 <= kk < Nz:
                        ref[i, j, k] += x[(i + dx) % Nx, (j + dy) % Ny, kk] @ w[t]
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
    # the adjoint: <conv(x), g> = <x, conv_transpose(g)>
    g = rng.normal(size=y.shape)
    assert np.sum(y * g) == pytest.approx(np.sum(x * L.conv_transpose(g, w, np.float64)), rel=1e-12)


def test_gradients_against_finite_differences():
    """Every layer's weight gradient (all seven, on all four levels) and the input gradient of the
    four-level network against central differences of the loss in FP64."""
    rng = np.random.default_rng(2)
    N = (128, 8, 8)
    x3 = rng.normal(size=N + (3,))
    x3[..., 2] = (rng.random(N) > 0.2)
    w = L.init_weights(3, scale=1.0).astype(np.float64)
    t = rng.normal(size=N)
    m = x3[..., 2].copy()
    out, st = L.forward(x3, w, precise=True)
    loss, dout = L.loss_and_grad_output(out, t, m)
    dw, gin = L.backward(st, w, dout, precise=True)

    def lossof(ww, xx):
        return L.loss_and_grad_output(L.forward(xx, ww, True)[0], t, m)[0]

    e = 1e-6
    for idx in [(0, 13, 0, 5), (0, 2, 2, 9), (1, 4, 7, 2), (2, 26, 15, 15), (3, 13, 9, 4), (4, 0, 3, 0),
                (5, 22, 11, 6), (6, 0, 3, 0), (6, 13, 9, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += e
        wm[idx] -= e
        fd = (lossof(wp, x3) - lossof(wm, x3)) / (2 * e)
        assert dw[idx] == pytest.approx(fd, rel=1e-6, abs=1e-10), idx
    for v in [(1, 2, 3, 0), (0, 0, 0, 1), (127, 7, 7, 2), (64, 4, 5, 0)]:
        xp, xm = x3.copy(), x3.copy()
        xp[v] += e
        xm[v] -= e
        fd = (lossof(w, xp) - lossof(w, xm)) / (2 * e)
        assert gin[v] == pytest.approx(fd, rel=1e-6, abs=1e-10), v


def test_level_transfers_against_brute_force():
    """pool is the mean of the 2×2×2 children, up copies the parent, and child_sum is up's adjoint:
    <up(c), f> = <c, child_sum(f)>."""
    rng = np.random.default_rng(5)
    f = rng.normal(size=(8, 6, 4, 3))
    p = L.pool(f, np.float64)
    for i, j, k in [(0, 0, 0), (3, 2, 1), (2, 1, 0)]:
        assert p[i, j, k] == pytest.approx(f[2 * i:2 * i + 2, 2 * j:2 * j + 2, 2 * k:2 * k + 2].mean(axis=(0, 1, 2)), rel=1e-14)
    c = rng.normal(size=(4, 3, 2, 3))
    u = L.up(c)
    assert u.shape == f.shape and u[5, 3, 2, 1] == c[2, 1, 1, 1]
    assert np.sum(u * f) == pytest.approx(np.sum(c * L.child_sum(f, np.float64)), rel=1e-12)


def test_network_sizes_are_checked():
    with pytest.raises(ValueError, match="multiple of 128"):
        L.forward(np.zeros((64, 8, 8, 3)), L.init_weights(0))


def test_first_adam_step_closed_form():
    """Step 1: m̂ = g, v̂ = g², so p ← p − lr · g / (|g| + ε)."""
    p = np.array([1.0, -2.0, 0.5], dtype=np.float32)
    g = np.array([0.3, -4.0, 1e-3], dtype=np.float32)
    z = np.zeros(3, dtype=np.float32)
    new, m, v = L.adam(p, g, z, z, 1, 0.01)
    np.testing.assert_allclose(new, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-6)


def _seamount_geometry():
    import oceananigans_mi355x as om
    from cases import seamount
    from oracle import oracle as O
    grid, kw, _ = seamount(N=(12, 10, 8))
    cfg, keep = om.build_config(grid, advection=om.WENO(),
                                pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=None, reltol=1e-14, abstol=0))
    o = O.OracleModel(cfg, grid)
    dzc, dzf, zc, zf = o.grid_z()
    H, Nz = grid.H[2], grid.Nz
    active = ~(zc[H:H + Nz][None, None, :] <= grid.bottom_heights()[:, :, None])
    dx, dy = grid.extent[0] / grid.Nx, grid.extent[1] / grid.Ny
    return grid, o, active, dx, dy, dzc[H:H + Nz], dzf[H:H + Nz]


def test_energy_gradient_matches_finite_differences():
    grid, o, active, dx, dy, dzc, dzf = _seamount_geometry()
    rng = np.random.default_rng(4)
    d = rng.normal(size=grid.N)
    E, g = L.energy(d, dx, dy, dzc, dzf, active)
    assert E > 0
    for _ in range(12):
        idx = tuple(rng.integers(n) for n in grid.N)
        h = 1e-6
        dp, dm = d.copy(), d.copy()
        dp[idx] += h
        dm[idx] -= h
        fd = (L.energy(dp, dx, dy, dzc, dzf, active)[0] - L.energy(dm, dx, dy, dzc, dzf, active)[0]) / (2 * h)
        assert fd == pytest.approx(g[idx], rel=1e-6, abs=1e-9 * abs(E))
    # constants have zero energy (the Neumann null space), inactive cells take no gradient
    assert L.energy(np.full(grid.N, 3.0), dx, dy, dzc, dzf, active)[0] == pytest.approx(0, abs=1e-9)
    assert np.all(g[~active] == 0)


def test_energy_is_the_cg_operators_quadratic_form():
    """For x solving the oracle's conditional CG problem V∇²x = V·rhs (immersed cells held at 0),
    xᵀ(−V∇²)x = −xᵀV·rhs: the energy of the restatement equals the form of the operator the
    oracle's CG inverts."""
    grid, o, active, dx, dy, dzc, dzf = _seamount_geometry()
    rng = np.random.default_rng(8)
    rhs = rng.normal(size=grid.N)
    rhs[~active] = 0
    V = dx * dy * np.broadcast_to(dzc[None, None, :], grid.N)
    rhs[active] -= (V[active] * rhs[active]).sum() / V[active].sum()
    x = o.poisson_solve(rhs)
    E, _ = L.energy(x, dx, dy, dzc, dzf, active)
    want = -float((x * V * rhs)[active].sum())
    assert E == pytest.approx(want, rel=1e-10)

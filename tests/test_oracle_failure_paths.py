"""Failure handling of the restated solvers, forced from the host on the CPU.

  conjugate_gradient_solver.jl:279-284   CG stops at maxiter (or at the tolerance), silently
  batched_tridiagonal_solver.jl:232-236  Thomas sweep: a pivot |β| ≤ 10 eps elides the update
"""
import numpy as np

import oceananigans_mi355x as om
from cases import convection, oracle_for
from oracle import oracle as O


def test_cg_stops_at_maxiter_above_tolerance():
    """With maxiter = 2 and an unreachable tolerance the solve stops after two iterations with the
    residual above the tolerance; with the default maxiter it converges."""
    from oceananigans_mi355x.synthetic import uniform
    grid = om.RectilinearGrid(size=(8, 6, 5), x=(0, 2.0), y=(0, 1.0), z=(-1.0, 0))
    rhs = uniform(3, grid.N, -1, 1)
    rhs -= rhs.mean()
    res = {}
    for maxiter in (2, None):
        kw = dict(advection=om.WENO(), pressure_solver=om.ConjugateGradientPoissonSolver(
            preconditioner=None, reltol=1e-13, abstol=0, maxiter=maxiter))
        o = oracle_for(grid, kw, {})
        o.poisson_solve(rhs)
        res[maxiter] = o.last_solve_info()
    it, r, _ = res[2]
    assert it == 2 and r > 1e-6
    it, r, true_r = res[None]
    assert it > 2 and true_r <= 1e-11 * np.linalg.norm(rhs)


def _thomas_reference_loop(a, b, c, f):
    """The reference's sweep (batched_tridiagonal_solver.jl:217-243) on one column, ϕ starting at 0."""
    n = len(b)
    phi, t = np.zeros(n), np.zeros(n)
    beta = b[0]
    phi[0] = f[0] / beta
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(1, n):
            t[k] = c[k - 1] / beta
            beta = b[k] - a[k - 1] * t[k]
            star = (f[k] - a[k - 1] * phi[k - 1]) / beta
            phi[k] = star if abs(beta) > 10 * np.finfo(float).eps else phi[k]
        for k in range(n - 2, -1, -1):
            phi[k] -= t[k + 1] * phi[k + 1]
    return phi


def test_thomas_pivot_guard_elides_the_update():
    """A column whose second pivot is exactly zero and one whose pivot falls below 10 eps: the guard
    keeps the output's previous value (0 here) at that level instead of dividing by the pivot, and
    the rest of the sweep follows the reference's loop; a regular column is solved exactly."""
    n = 5
    a = np.array([[1.0, 1.0, 1.0, 1.0], [1.0, 1.0, 1.0, 1.0], [0.5, 0.5, 0.5, 0.5]])
    c = np.array([[1.0, 1.0, 1.0, 1.0], [1.0, 1.0, 1.0, 1.0], [0.5, 0.5, 0.5, 0.5]])
    b = np.array([[1.0, 1.0, 3.0, 3.0, 3.0],                   # β₂ = 1 − 1·1 = 0 exactly
                  [1.0, 1.0 + 4e-16, 3.0, 3.0, 3.0],           # β₂ = 4.4e-16 ≤ 10 eps
                  [2.0, 2.0, 2.0, 2.0, 2.0]])                  # diagonally dominant
    f = np.array([[1.0, 2.0, 3.0, 4.0, 5.0]] * 3)
    x = O.thomas(a, b, c, f)
    for s in range(2):
        want = _thomas_reference_loop(a[s], b[s], c[s], f[s])
        np.testing.assert_array_equal(np.isfinite(x[s]), np.isfinite(want))
        fin = np.isfinite(want)
        np.testing.assert_allclose(x[s][fin], want[fin], rtol=1e-15, atol=0)
    # the guard fired at level 2: without it the division by the tiny pivot dominates
    assert abs(x[1][1]) < 1e3
    A = np.diag(b[2]) + np.diag(a[2], -1) + np.diag(c[2], 1)
    np.testing.assert_allclose(A @ x[2], f[2], rtol=1e-14)


def test_nonfinite_state_is_contained_to_nan():
    """A NaN planted in the state stays a NaN (no exception, no trap): the oracle steps on, and the
    NaN reaches the fields, which is what the device NaN check looks for."""
    grid, kw, init = convection(N=(8, 8, 8), noise_velocity=1e-2)
    o = oracle_for(grid, kw, init)
    u = o.get(0)
    u[3, 3, 3] = np.nan
    o.set(0, u)
    o.time_step(1.0, 1)
    assert np.isnan(o.get(0)).any()

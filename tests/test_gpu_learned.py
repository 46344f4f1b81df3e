"""The learned preconditioner on the GPU, through the C ABI: the MFMA forward pass and the
hand-written backward pass against the oracle (oracle/learned.py), training on pairs logged from
the model's own CG solves, and flexible CG with the trained network against FFT-preconditioned CG.

Forward / backward bars: relative 2-norm difference ≤ 2e-3 per output.  Both sides round the
same quantities to BF16 and accumulate in FP32, in different orders; a sum that lands within
FP32 rounding of a BF16 rounding boundary rounds the other way (one BF16 ulp, 0.4 %), which is
what the bar allows for."""
import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import product_for, rel_err, seamount
from oracle import learned as L

pytestmark = pytest.mark.gpu


def _net_model(N=(128, 8, 8)):
    grid = om.RectilinearGrid(size=N, extent=(1.0, 0.5, 0.4))
    return om.NonhydrostaticModel(grid, advection=om.WENO(),
                                  pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=om.LearnedPreconditioner()))


def _inputs(N, seed):
    rng = np.random.default_rng(seed)
    x3 = rng.normal(size=N + (3,)).astype(np.float32)
    x3[..., 2] = (rng.random(N) > 0.15).astype(np.float32)
    return x3


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_matches_oracle():
    model = _net_model()
    N = model.grid.N
    w = L.init_weights(5, scale=1.0)
    model.learned_set_weights(w)
    np.testing.assert_array_equal(model.learned_get_weights(), w)
    x3 = _inputs(N, 1)
    y = model.learned_forward(x3)
    yo, _ = L.forward(x3, w)
    e = _rel(y, yo)
    print("forward rel err", e, "max|y|", np.abs(yo).max(), flush=True)
    assert e <= 2e-3


def test_backward_matches_oracle():
    model = _net_model()
    N = model.grid.N
    w = L.init_weights(6, scale=1.0)
    model.learned_set_weights(w)
    x3 = _inputs(N, 2)
    t = np.random.default_rng(3).normal(size=N).astype(np.float32)
    m = x3[..., 2]
    loss, dw, din = model.learned_backward(x3, t, m)
    out, acts = L.forward(x3, w)
    loss_o, dout = L.loss_and_grad_output(out, t, m)
    dwo, gino = L.backward(acts, w, dout.astype(np.float32))
    errs = [_rel(dw[l], dwo[l]) for l in range(L.NL)]
    ein = _rel(din, gino[..., :3])
    print("loss", loss, loss_o, "dW rel err per layer", errs, "input grad", ein, flush=True)
    assert loss == pytest.approx(loss_o, rel=2e-3)
    assert max(errs) <= 2e-3 and ein <= 2e-3


@pytest.mark.parametrize("N", [(128, 8, 8), (256, 24, 16)])
def test_tiled_and_per_tap_convolutions_agree_bitwise(N):
    """The convolutions with LDS-staged input tiles and the ones that read every tap from memory use
    the same operands in the same MFMA order: forward outputs and every gradient agree bit for bit
    (Ny = 24, Nz = 16 leave partly filled tiles on the coarse levels)."""
    model = _net_model(N)
    model.learned_set_weights(L.init_weights(9, scale=1.0))
    x3 = _inputs(N, 4)
    t = np.random.default_rng(5).normal(size=N).astype(np.float32)
    res = {}
    for tiled in (True, False):
        model.set_tuning(conv_tiled=tiled)
        res[tiled] = (model.learned_forward(x3),) + tuple(model.learned_backward(x3, t, x3[..., 2]))
    for a, b in zip(res[True], res[False]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
    model.close()


def test_matrix_core_weight_gradients_match_the_fp32_sums():
    """The weight gradients on the matrix cores (the default) and as FP32 sums: the same BF16
    products summed in different orders, equal to FP32 rounding; both against the oracle."""
    N = (256, 16, 16)
    model = _net_model(N)
    w = L.init_weights(12, scale=1.0)
    model.learned_set_weights(w)
    x3 = _inputs(N, 6)
    t = np.random.default_rng(7).normal(size=N).astype(np.float32)
    grads = {}
    for mfma in (True, False):
        model.set_tuning(wgrad_mfma=mfma)
        grads[mfma] = model.learned_backward(x3, t, x3[..., 2])[1]
    out, st = L.forward(x3, w)
    _, dout = L.loss_and_grad_output(out, t, x3[..., 2])
    dwo, _ = L.backward(st, w, dout.astype(np.float32))
    between = [_rel(grads[True][l], grads[False][l]) for l in range(L.NL)]
    vs_oracle = [_rel(grads[True][l], dwo[l]) for l in range(L.NL)]
    print("MFMA vs FP32-sum weight gradients per layer", between, "MFMA vs oracle", vs_oracle, flush=True)
    assert max(between) <= 1e-5
    assert max(vs_oracle) <= 2e-3
    model.close()


def _seamount_models(pc_a, pc_b, N=(128, 32, 16)):
    grid, kw, init = seamount(N=N)
    ka = dict(kw, pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=pc_a, reltol=1e-10, abstol=0))
    kb = dict(kw, pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=pc_b, reltol=1e-10, abstol=0))
    return product_for(grid, ka, init), product_for(grid, kb, init)


def test_training_gain_on_held_out_solves_and_flexible_cg_matches_fft_cg():
    """Pairs logged from the learned model's own solves (zero output layer: it starts as the FFT
    preconditioner); four rounds of one time step of logging and 400 Adam steps on the A-energy
    loss.  The gain is measured on pairs from later solves
    that training never saw: the relative A-energy of the error the network leaves,
    (δ − t)ᵀ(−V∇²)(δ − t) / tᵀ(−V∇²)t, against 1 for the FFT preconditioner alone.  Flexible CG
    with the trained network reaches the same states as FFT-preconditioned CG and the recomputed
    FP64 residual stays within a hundred times the recursive one; the iteration counts are reported."""
    learned, fft = _seamount_models(om.LearnedPreconditioner(), om.FourierTridiagonalPoissonSolver())
    w0 = L.init_weights(7)
    w0[L.NL - 1] = 0.0
    learned.learned_set_weights(w0)
    # rounds of logging and training, so that the pairs include residuals of the network's own solves
    learned.learned_log(iterations_per_solve=8, capacity=96)
    before = None
    for _ in range(4):
        om.time_step_(learned, 20.0)
        om.time_step_(fft, 20.0)
        if before is None:
            before = learned.learned_evaluate()
        learned.learned_train(400, lr=5e-4)
    npairs = learned.learned_pairs()
    train_after = learned.learned_evaluate()
    learned.learned_log(iterations_per_solve=8, capacity=48)        # a new set: held-out pairs
    its = []
    for _ in range(2):
        om.time_step_(learned, 20.0)
        om.time_step_(fft, 20.0)
        its.append((learned.last_solve_info()[0], fft.last_solve_info()[0], learned.learned_last_solve()))
    held = learned.learned_evaluate()
    print("pairs", npairs, "energy loss on training pairs before / after", before, train_after,
          "on", learned.learned_pairs(), "held-out pairs", held, flush=True)
    print("iterations learned / fft / (network iterations, fell back)", its, flush=True)
    assert before == pytest.approx(1.0, abs=1e-6)     # zero correction = the FFT preconditioner
    assert train_after < 0.6 and held < 0.8
    for n in ("u", "v", "w"):
        assert rel_err(learned.velocities[n].interior(), fft.velocities[n].interior()) <= 1e-8
    it, res, true_res = learned.last_solve_info()
    assert true_res <= 100 * res + 1e-30 or true_res <= 1e-8


def test_network_input_written_directly_equals_the_dense_input_path():
    """In flexible CG the network input is written straight into its padded BF16 box, except in the
    iterations that log a pair, which keep the dense FP32 input and convert it.  Two models step from
    the same state with the same weights, one logging its first 8 iterations per solve and one not:
    the same channels reach the network either way, so the states are bit-identical."""
    w = L.init_weights(5)
    a, b = _seamount_models(om.LearnedPreconditioner(weights=w), om.LearnedPreconditioner(weights=w), N=(128, 16, 16))
    a.learned_log(iterations_per_solve=8, capacity=32)
    for _ in range(2):
        om.time_step_(a, 20.0)
        om.time_step_(b, 20.0)
    print("iterations logged / direct", a.last_solve_info(), b.last_solve_info(), a.learned_last_solve(),
          b.learned_last_solve(), flush=True)
    for n in ("u", "v", "w"):
        assert np.array_equal(a.velocities[n].interior(), b.velocities[n].interior()), n
    assert np.array_equal(a.pressures.pNHS.interior(), b.pressures.pNHS.interior())
    a.close()
    b.close()


def test_untrained_large_weights_hand_over_to_fft():
    """A network far from useful (large random weights, output layer scaled up): flexible CG stalls,
    hands over to the FFT preconditioner (three iterations without progress, or the recomputed FP64
    residual above a hundred times the tolerance), and the state still matches FFT-preconditioned CG."""
    learned, fft = _seamount_models(om.LearnedPreconditioner(weights=L.init_weights(11, scale=50.0)),
                                    om.FourierTridiagonalPoissonSolver(), N=(128, 16, 16))
    for _ in range(2):
        om.time_step_(learned, 20.0)
        om.time_step_(fft, 20.0)
        its, fell_back = learned.learned_last_solve()
        print("learned", learned.last_solve_info(), (its, fell_back), "fft", fft.last_solve_info(), flush=True)
        assert fell_back and its >= 1
    for n in ("u", "v", "w"):
        assert rel_err(learned.velocities[n].interior(), fft.velocities[n].interior()) <= 1e-8


def test_energy_loss_and_gradient_match_oracle():
    """The loss the training back-propagates: E = dᵀ(−V∇²)d, d = active·(y − t), with the CG
    solver's conditional operator on the immersed seamount, and dE/dy, against the NumPy
    restatement (oracle/learned.py energy, itself pinned to the oracle CG operator's quadratic
    form and to finite differences)."""
    grid, kw, init = seamount(N=(128, 16, 16))
    kw = dict(kw, pressure_solver=om.ConjugateGradientPoissonSolver(preconditioner=om.LearnedPreconditioner()))
    model = om.NonhydrostaticModel(grid, **kw)
    from oracle import oracle as O
    cfg, keep = om.build_config(grid, **kw)
    dzc, dzf, zc, zf = O.OracleModel(cfg, grid).grid_z()
    H, Nz = grid.H[2], grid.Nz
    active = ~(zc[H:H + Nz][None, None, :] <= grid.bottom_heights()[:, :, None])
    rng = np.random.default_rng(21)
    y = rng.normal(size=grid.N).astype(np.float32)
    t = rng.normal(size=grid.N).astype(np.float32)
    E, g = model.learned_energy_loss(y, t)
    dx, dy = grid.extent[0] / grid.Nx, grid.extent[1] / grid.Ny
    Eo, go = L.energy(y.astype(np.float64) - t.astype(np.float64), dx, dy, dzc[H:H + Nz], dzf[H:H + Nz], active)
    e = _rel(g, go)
    print("energy", E, Eo, "gradient rel err", e, flush=True)
    assert E == pytest.approx(Eo, rel=1e-12)
    assert e <= 1e-6
    model.close()

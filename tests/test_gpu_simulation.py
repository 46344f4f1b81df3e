"""Device diagnostics and the Simulation layer on the GPU, through the C ABI:
cell_advection_timescale against the oracle (exact: max and min are exact operations and the
per-cell expression is evaluated in the reference's order), field maxima / NaN detection against
the parent data read back, the TimeStepWizard and NaNChecker in run!, and checkpoint / pickup
continuing bit for bit."""
import math

import numpy as np
import pytest

import oceananigans_mi355x as om
from cases import convection, product_for, rel_err, seamount

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("float_type", [np.float64, np.float32])
def test_cell_advection_timescale_matches_oracle(float_type):
    grid, kw, init = convection(N=(24, 20, 16), noise_velocity=1e-2, float_type=float_type)
    model = product_for(grid, kw, init)
    om.time_step_(model, 1.0)
    tau = model.cell_advection_timescale()
    from oracle.oracle import OracleModel
    o = OracleModel(om.build_config(grid, **kw)[0], grid)          # the oracle in the model's precision
    for q, n in enumerate(("u", "v", "w")):
        o.set(q, model.velocities[n].interior())
    tau_o = o.cell_advection_timescale()
    print(float_type.__name__, "tau gpu / oracle", tau, tau_o, flush=True)
    assert tau == tau_o
    assert math.isfinite(tau) and tau > 0


def test_field_statistics_and_nan_detection():
    grid, kw, init = convection(N=(16, 16, 8), noise_velocity=1e-2)
    model = product_for(grid, kw, init)
    om.time_step_(model, 1.0)
    for f in (model.velocities.u, model.velocities.w, model.tracers.b, model.closure_fields.nu):
        p = f.parent()
        mx, mn, nn = model.field_statistics(f.id, True)
        assert (mx, mn, nn) == (p.max(), p.min(), 0)
        i = f.interior()
        assert model.field_statistics(f.id, False)[:2] == (i.max(), i.min())
    p = model.velocities.u.parent()
    p[0, 0, 0] = np.nan                                   # a halo point: hasnan looks at the parent
    model.velocities.u.set_parent(p)
    assert model.velocities.u.hasnan()
    assert model.field_statistics(model.velocities.u.id, False)[2] == 0


def test_nan_checker_stops_the_simulation():
    grid, kw, init = convection(N=(16, 16, 8))
    model = product_for(grid, kw, init)
    sim = om.Simulation(model, Δt=1.0, stop_iteration=50)
    sim.callbacks["nan_checker"].schedule = om.IterationInterval(1)
    p = model.velocities.u.parent()
    p[5, 5, 5] = np.nan
    model.velocities.u.set_parent(p)
    om.run_(sim)
    assert sim.callbacks["nan_checker"].func.nan_detected
    assert model.clock.iteration < 50


def test_time_step_wizard_uses_the_device_timescale():
    grid, kw, init = convection(N=(16, 16, 8), noise_velocity=5e-2)
    model = product_for(grid, kw, init)
    sim = om.Simulation(model, Δt=0.01, stop_iteration=4)
    om.conjure_time_step_wizard_(sim, om.IterationInterval(1), cfl=0.3, max_change=100.0, min_change=0.01)
    seen = []
    sim.callbacks["record"] = om.Callback(lambda s: seen.append((s.Δt, s.model.cell_advection_timescale())))
    om.run_(sim)
    for dt, tau in seen[1:]:
        assert dt == pytest.approx(0.3 * tau, rel=1e-12)


@pytest.mark.parametrize("case", ["convection_fp32", "seamount_cg_fp64"])
def test_checkpoint_pickup_continues_bit_for_bit(case, tmp_path):
    """Run 6 steps straight, and 3 steps + checkpoint + pickup in a fresh model + 3 steps: the
    states agree exactly (the CG warm start is part of the saved state)."""
    if case == "convection_fp32":
        grid, kw, init = convection(N=(16, 16, 16), noise_velocity=1e-2, float_type=np.float32)
        dt = 1.0
    else:
        grid, kw, init = seamount(N=(16, 16, 12))
        dt = 20.0
    a = product_for(grid, kw, init)
    sim = om.Simulation(a, Δt=dt, stop_iteration=3)
    sim.output_writers["checkpointer"] = om.Checkpointer(a, om.IterationInterval(3), dir=str(tmp_path), prefix="cp")
    om.run_(sim)
    sim.stop_iteration = 6
    om.run_(sim)
    b = om.NonhydrostaticModel(grid, **kw)
    simb = om.Simulation(b, Δt=dt, stop_iteration=6)
    simb.output_writers["checkpointer"] = om.Checkpointer(b, om.IterationInterval(1000), dir=str(tmp_path), prefix="cp")
    om.run_(simb, pickup=3)
    assert b.clock.iteration == 6 and b.clock.time == a.clock.time
    for n in ("u", "v", "w"):
        np.testing.assert_array_equal(b.velocities[n].interior(), a.velocities[n].interior())
    for n in a.tracers:
        np.testing.assert_array_equal(b.tracers[n].interior(), a.tracers[n].interior())
    np.testing.assert_array_equal(b.pressures.pNHS.interior(), a.pressures.pNHS.interior())


@pytest.mark.parametrize("case", ["convection_fp64", "seamount_cg_fp64"])
def test_gpu_checkpoint_resumes_in_the_oracle(case, tmp_path):
    """A checkpoint written by the GPU model (parent arrays under the reference's JLD2 group names,
    checkpointer.jl:7-165, nonhydrostatic_model.jl:354-375) is read by an independent reader, the
    oracle, which continues from it; its state after 3 more steps equals the GPU model's."""
    from cases import oracle_for
    from oceananigans_mi355x.simulations import load_checkpoint_state
    if case == "convection_fp64":
        grid, kw, init = convection(N=(16, 16, 16), noise_velocity=1e-2)
        dt, tol = 1.0, 1e-11
    else:
        grid, kw, init = seamount(N=(16, 16, 12))
        dt, tol = 20.0, 1e-10
    a = product_for(grid, kw, init)
    sim = om.Simulation(a, Δt=dt, stop_iteration=3)
    sim.output_writers["checkpointer"] = om.Checkpointer(a, om.IterationInterval(3), dir=str(tmp_path), prefix="cp")
    om.run_(sim)
    state = load_checkpoint_state(str(tmp_path / "cp_iteration3.npz"))
    orc = oracle_for(grid, kw, {})
    names = ("u", "v", "w") + tuple(kw["tracers"])
    orc.load_state(state, names)
    assert orc.clock()[0] == a.clock.time and orc.clock()[1] == 3
    for _ in range(3):
        om.time_step_(a, dt)
    orc.time_step(dt, 3)
    errs = {n: rel_err((a.velocities[n] if q < 3 else a.tracers[n]).interior(), orc.get(q)) for q, n in enumerate(names)}
    print(case, "GPU checkpoint → oracle, 3 more steps", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= tol
    assert orc.clock()[0] == pytest.approx(a.clock.time, rel=1e-14) and orc.clock()[1] == 6
    a.close()

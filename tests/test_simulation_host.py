"""Host logic of Simulation / run! / schedules / TimeStepWizard (src/Simulations/, src/Utils/schedules.jl)
on a stand-in model that only keeps a clock: no GPU and no library calls."""
import math

import pytest

import oceananigans_mi355x as om
from oceananigans_mi355x.simulations import Simulation, TimeInterval, IterationInterval, Callback


class _Clock:
    def __init__(self):
        self.time, self.iteration, self.stage = 0.0, 0, 1

    def _get(self):
        return self.time, self.iteration, self.stage, 0.0, 0.0


class _Field:
    def hasnan(self):
        return False


class _Model:
    """Advances its clock only; the advective timescale is a settable number."""

    def __init__(self, tau=1.0):
        self.clock = _Clock()
        self.velocities = {"u": _Field()}
        self.tau = tau
        self.steps = []

    def update_state(self):
        pass

    def time_step(self, dt, n=1):
        self.steps.append(dt)
        self.clock.time += dt
        self.clock.iteration += 1

    def cell_advection_timescale(self):
        return self.tau


def test_wizard_limits_and_arguments():
    """new_time_step: cfl·τ clamped by max_change/min_change of the old Δt and by [min_Δt, max_Δt]
    (time_step_wizard.jl:97-111); invalid change factors raise (:73-74)."""
    m = _Model(tau=10.0)
    w = om.TimeStepWizard(cfl=0.5, max_change=1.1, min_change=0.5)
    assert w.new_time_step(1.0, m) == pytest.approx(1.1)        # 5.0 capped at 1.1·Δt
    m.tau = 0.1
    assert w.new_time_step(1.0, m) == pytest.approx(0.5)        # 0.05 raised to 0.5·Δt
    m.tau = 3.0
    assert w.new_time_step(1.4, m) == pytest.approx(1.5)        # inside the band: cfl·τ
    w2 = om.TimeStepWizard(cfl=0.5, max_Δt=0.2, min_change=0.1)
    assert w2.new_time_step(1.0, m) == pytest.approx(0.2)
    with pytest.raises(ValueError):
        om.TimeStepWizard(min_change=1.0)
    with pytest.raises(ValueError):
        om.TimeStepWizard(max_change=1.0)


def test_stop_criteria_and_stop_time_alignment():
    """run! stops on stop_iteration / stop_time; the last step is shortened to land on stop_time
    (run.jl:33-48)."""
    m = _Model()
    sim = Simulation(m, Δt=0.3, stop_time=1.0)
    om.run_(sim)
    assert m.steps[:3] == [0.3, 0.3, 0.3]
    assert m.steps[-1] == pytest.approx(0.1)
    assert m.clock.time == pytest.approx(1.0)
    m2 = _Model()
    om.run_(Simulation(m2, Δt=0.1, stop_iteration=7))
    assert m2.clock.iteration == 7


def test_time_interval_callbacks_align_time_steps():
    """A TimeInterval(0.25) callback shortens steps to land on its actuation times and fires
    once per interval (schedules.jl:81-114); IterationInterval fires on multiples of the interval."""
    m = _Model()
    sim = Simulation(m, Δt=0.1, stop_time=1.0)
    fired, every3 = [], []
    sim.callbacks["t"] = Callback(lambda s: fired.append(round(s.model.clock.time, 12)), TimeInterval(0.25))
    sim.callbacks["i"] = Callback(lambda s: every3.append(s.model.clock.iteration), IterationInterval(3))
    om.run_(sim)
    assert fired == [0.0, 0.25, 0.5, 0.75, 1.0]      # initialization call at t = 0, then each interval
    assert all(i % 3 == 0 for i in every3) and every3[0] == 0


def test_wizard_as_callback_updates_the_simulation_step():
    m = _Model(tau=2.0)
    sim = Simulation(m, Δt=0.5, stop_iteration=6)
    om.conjure_time_step_wizard_(sim, IterationInterval(1), cfl=0.5, max_change=1.5)
    om.run_(sim)
    # 0.5 → min(1.5·0.5, 0.5·2.0) = 0.75 → min(1.125, 1.0) = 1.0 → stays 1.0
    assert m.steps[0] == pytest.approx(0.75) and m.steps[1:] == pytest.approx([1.0] * 5)

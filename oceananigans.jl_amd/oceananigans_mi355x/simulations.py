"""Simulation and run! (src/Simulations/simulation.jl:67-81, src/Simulations/run.jl:151-269).

Only the time-stepping loop, stop criteria and callbacks are mirrored; output writers and the
TimeStepWizard are out of scope.
"""
import time as _time


class Callback:
    def __init__(self, func, schedule_iterations=1):
        self.func, self.every = func, int(schedule_iterations)


class Simulation:
    def __init__(self, model, Δt=None, stop_iteration=None, stop_time=None, wall_time_limit=None, **kw):
        if Δt is None:
            Δt = kw.pop("dt", None)
        if Δt is None:
            raise ValueError("Δt is required")
        self.model, self.Δt = model, float(Δt)
        self.stop_iteration, self.stop_time, self.wall_time_limit = stop_iteration, stop_time, wall_time_limit
        self.callbacks = {}
        self.running = False
        self.run_wall_time = 0.0

    def add_callback(self, name, func, iterations=1):
        self.callbacks[name] = Callback(func, iterations)

    def _done(self, t0):
        clock = self.model.clock
        if self.stop_iteration is not None and clock.iteration >= self.stop_iteration:
            return True
        if self.stop_time is not None and clock.time >= self.stop_time - 1e-12 * max(1.0, abs(self.stop_time)):
            return True
        if self.wall_time_limit is not None and _time.time() - t0 >= self.wall_time_limit:
            return True
        return False


def run_(sim):
    """run!(simulation): time_step! until a stop criterion is met (run.jl:151-269)."""
    sim.running = True
    t0 = _time.time()
    while not sim._done(t0):
        dt = sim.Δt
        if sim.stop_time is not None:
            dt = min(dt, sim.stop_time - sim.model.clock.time)   # align with stop_time
        sim.model.time_step(dt, 1)
        it = sim.model.clock.iteration
        for cb in sim.callbacks.values():
            if it % cb.every == 0:
                cb.func(sim)
    sim.running = False
    sim.run_wall_time += _time.time() - t0

"""Simulation, run!, schedules, callbacks, TimeStepWizard, NaNChecker and Checkpointer
(src/Simulations/simulation.jl, run.jl, callback.jl, time_step_wizard.jl; src/Utils/schedules.jl;
src/Diagnostics/nan_checker.jl; src/OutputWriters/checkpointer.jl).

The loop, stop criteria, schedules and time-step alignment follow the reference.  The reductions
they need (the advective timescale, field maxima, NaN detection) run on the device through the C
ABI; only scalars cross to the host.  Checkpoints hold the reference's prognostic state
(clock, velocities, tracers, Gⁿ/G⁻, closure fields) with parent arrays (halos included) under the
reference's JLD2 group names, written as a NumPy ``.npz`` archive; the non-hydrostatic pressure is
stored too, so that a conjugate-gradient solve resumes from the same warm start.
"""
import glob
import math
import os
import re
import time as _time

import numpy as np


# ---------------------------------------------------------------------------------------------
# schedules (src/Utils/schedules.jl)
# ---------------------------------------------------------------------------------------------
class IterationInterval:
    """Actuates when (iteration − offset) is a multiple of `interval` (schedules.jl:136-160)."""

    def __init__(self, interval, offset=0):
        self.interval, self.offset = int(interval), int(offset)

    def initialize(self, model):
        pass

    def __call__(self, model):
        return (model.clock.iteration - self.offset) % self.interval == 0

    def aligned_time_step(self, clock, dt):
        return dt


class TimeInterval:
    """Actuates every `interval` of simulation time from the time it was initialized
    (schedules.jl:32-114); time steps are shortened to land on actuation times."""

    def __init__(self, interval):
        self.interval = float(interval)
        self.first_actuation_time = 0.0
        self.actuations = 0

    def initialize(self, model):
        self.first_actuation_time = model.clock.time
        self.actuations = 0

    def next_actuation_time(self):
        return self.first_actuation_time + (self.actuations + 1) * self.interval

    def __call__(self, model):
        t, ts = model.clock.time, self.next_actuation_time()
        if t >= ts:
            self.actuations += 1
            # advance so that the next actuation lies strictly in the future
            while self.next_actuation_time() <= t:
                self.actuations += 1
            return True
        return False

    def aligned_time_step(self, clock, dt):
        return min(dt, self.next_actuation_time() - clock.time)


class Callback:
    """Callback(func, schedule=IterationInterval(1); parameters) (src/Simulations/callback.jl)."""

    def __init__(self, func, schedule=None, parameters=None):
        if isinstance(schedule, int):
            schedule = IterationInterval(schedule)
        self.func, self.schedule, self.parameters = func, schedule or IterationInterval(1), parameters

    def __call__(self, sim):
        if self.parameters is None:
            return self.func(sim)
        return self.func(sim, self.parameters)


# ---------------------------------------------------------------------------------------------
# diagnostics (src/Diagnostics/nan_checker.jl, src/Simulations/time_step_wizard.jl)
# ---------------------------------------------------------------------------------------------
class NaNChecker:
    """Stops the simulation (or raises, when `erroring`) if a field's parent data holds a NaN;
    the check is a device reduction (nan_checker.jl:37-63)."""

    def __init__(self, fields, erroring=False):
        self.fields, self.erroring, self.nan_detected = dict(fields), erroring, False

    def __call__(self, sim):
        for name, field in self.fields.items():
            if field.hasnan():
                sim.running = False
                self.nan_detected = True
                msg = f"time = {sim.model.clock.time}, iteration = {sim.model.clock.iteration}: NaN found in field {name}."
                if self.erroring:
                    raise FloatingPointError(msg + " Aborting simulation.")
                if sim.verbose:
                    print(msg + " Stopping simulation.")


def cell_advection_timescale(model):
    return model.cell_advection_timescale()


def _min_spacing(grid, horizontal=True, vertical=True):
    from .grids import Flat
    out = []
    if horizontal:
        for d in (0, 1):
            if grid.topology[d] is not Flat:
                out.append(grid.extent[d] / grid.N[d])
    if vertical and grid.topology[2] is not Flat:
        out.append(float(np.min(np.diff(grid.z_faces))) if grid.stretched else grid.extent[2] / grid.N[2])
    return min(out)


def cell_diffusion_timescale(model):
    """Δ² / max diffusivity for the model's closure (turbulence_closure_diagnostics.jl:44-70):
    the eddy viscosity and diffusivities are reduced over their parent arrays on the device."""
    from .physics import AnisotropicMinimumDissipation, ScalarDiffusivity, SmagorinskyLilly
    closure = model.closure
    if closure is None:
        return math.inf
    d = _min_spacing(model.grid)
    if isinstance(closure, ScalarDiffusivity):
        nu, kappa = closure.nu, closure.kappa
        return min(d ** 2 / nu if nu else math.inf, d ** 2 / kappa if kappa else math.inf)
    nu = model.closure_fields.nu.maximum()
    if isinstance(closure, SmagorinskyLilly):
        return d ** 2 / (nu * max(1.0, 1.0 / closure.Pr))
    if isinstance(closure, AnisotropicMinimumDissipation):
        kap = [f.maximum() for n, f in model.closure_fields.items() if n.startswith("kappa_")]
        return min(d ** 2 / nu, d ** 2 / max(kap)) if kap else d ** 2 / nu
    return math.inf


class TimeStepWizard:
    """Adjusts Δt toward the target advective (and diffusive) CFL numbers within
    [max(min_Δt, min_change·Δt), min(max_Δt, max_change·Δt)] (time_step_wizard.jl:64-115)."""

    def __init__(self, cfl=0.2, diffusive_cfl=math.inf, max_change=1.1, min_change=0.5, max_Δt=math.inf, min_Δt=0.0,
                 cell_advection_timescale=cell_advection_timescale, cell_diffusion_timescale=None):
        if min_change >= 1:
            raise ValueError(f"min_change must be < 1. You provided min_change = {min_change}.")
        if max_change <= 1:
            raise ValueError(f"max_change must be > 1. You provided max_change = {max_change}.")
        if cell_diffusion_timescale is None:
            cell_diffusion_timescale = globals()["cell_diffusion_timescale"] if math.isfinite(diffusive_cfl) else (lambda m: math.inf)
        self.cfl, self.diffusive_cfl = cfl, diffusive_cfl
        self.max_change, self.min_change, self.max_Δt, self.min_Δt = max_change, min_change, max_Δt, min_Δt
        self.cell_advection_timescale, self.cell_diffusion_timescale = cell_advection_timescale, cell_diffusion_timescale

    def new_time_step(self, old_dt, model):
        advective = self.cfl * self.cell_advection_timescale(model)
        diffusive = self.diffusive_cfl * self.cell_diffusion_timescale(model)
        dt = min(advective, diffusive)
        dt = min(self.max_change * old_dt, dt)
        dt = max(self.min_change * old_dt, dt)
        return min(max(dt, self.min_Δt), self.max_Δt)

    def __call__(self, sim):
        sim.Δt = self.new_time_step(sim.Δt, sim.model)


def conjure_time_step_wizard_(sim, schedule=None, **wizard_kw):
    """conjure_time_step_wizard!(simulation, schedule=IterationInterval(10); kw...)."""
    sim.callbacks["time_step_wizard"] = Callback(TimeStepWizard(**wizard_kw), schedule or IterationInterval(10))


# ---------------------------------------------------------------------------------------------
# prognostic state and Checkpointer (nonhydrostatic_model.jl:354-375, checkpointer.jl)
# ---------------------------------------------------------------------------------------------
def prognostic_state(model):
    """Host copy of the model's prognostic state: {group path: array}, parent data with halos."""
    t, it, stage, ldt, lsdt = model.clock._get()
    state = {"clock/time": np.float64(t), "clock/iteration": np.int64(it), "clock/stage": np.int64(stage),
             "clock/last_Δt": np.float64(ldt), "clock/last_stage_Δt": np.float64(lsdt)}
    for n, f in model.velocities.items():
        state[f"velocities/{n}/data"] = f.parent()
    for n, f in model.tracers.items():
        state[f"tracers/{n}/data"] = f.parent()
    for n in model.Gn:
        state[f"timestepper/Gⁿ/{n}/data"] = model.Gn[n].parent()
        state[f"timestepper/G⁻/{n}/data"] = model.Gm[n].parent()
    for n, f in model.closure_fields.items():
        state[f"closure_fields/{n}/data"] = f.parent()
    state["pressures/pNHS/data"] = model.pressures.pNHS.parent()
    return state


def restore_prognostic_state_(model, state):
    """restore_prognostic_state!(model, state): copy every saved parent array back to the device."""
    lib, h = model._lib, model._h
    from . import _abi
    _abi.check(lib.om_set_clock(h, float(state["clock/time"]), int(state["clock/iteration"]), int(state["clock/stage"]),
                                float(state["clock/last_Δt"]), float(state["clock/last_stage_Δt"])), h)
    groups = [("velocities", model.velocities), ("tracers", model.tracers), ("closure_fields", model.closure_fields)]
    for g, fields in groups:
        for n, f in fields.items():
            key = f"{g}/{n}/data"
            if key in state:
                f.set_parent(state[key])
    for n in model.Gn:
        for tag, fields in (("Gⁿ", model.Gn), ("G⁻", model.Gm)):
            key = f"timestepper/{tag}/{n}/data"
            if key in state:
                fields[n].set_parent(state[key])
    if "pressures/pNHS/data" in state:
        model.pressures.pNHS.set_parent(state["pressures/pNHS/data"])
    return model


class Checkpointer:
    """Checkpointer(model; schedule, dir=".", prefix="checkpoint", cleanup=False): writes
    `{prefix}_iteration{N}.npz` with the prognostic state (checkpointer.jl:7-165)."""

    def __init__(self, model, schedule, dir=".", prefix="checkpoint", cleanup=False):
        self.model, self.schedule, self.dir, self.prefix, self.cleanup = model, schedule, dir, prefix, cleanup
        os.makedirs(dir, exist_ok=True)

    def path(self, iteration):
        return os.path.join(self.dir, f"{self.prefix}_iteration{int(iteration)}.npz")

    def paths(self):
        pat = re.compile(re.escape(self.prefix) + r"_iteration(\d+)\.npz$")
        found = []
        for p in glob.glob(os.path.join(self.dir, self.prefix + "_iteration*.npz")):
            m = pat.search(os.path.basename(p))
            if m:
                found.append((int(m.group(1)), p))
        return sorted(found)

    def write(self, sim):
        it = self.model.clock.iteration
        if self.cleanup:
            for _, p in self.paths():
                os.remove(p)
        np.savez(self.path(it), **prognostic_state(self.model))


def load_checkpoint_state(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ---------------------------------------------------------------------------------------------
# Simulation and run! (simulation.jl, run.jl)
# ---------------------------------------------------------------------------------------------
def _stop_iteration_exceeded(sim):
    if sim.model.clock.iteration >= sim.stop_iteration:
        sim.running = False


def _stop_time_exceeded(sim):
    if sim.model.clock.time >= sim.stop_time:
        sim.running = False


def _wall_time_limit_exceeded(sim):
    if sim.run_wall_time >= sim.wall_time_limit:
        sim.running = False


class Simulation:
    def __init__(self, model, Δt=None, stop_iteration=math.inf, stop_time=math.inf, wall_time_limit=math.inf,
                 align_time_step=True, minimum_relative_step=0.0, verbose=False, **kw):
        if Δt is None:
            Δt = kw.pop("dt", None)
        if Δt is None:
            raise ValueError("Δt is required")
        if kw:
            raise TypeError(f"unexpected arguments {sorted(kw)}")
        self.model, self.Δt = model, float(Δt)
        self.stop_iteration = math.inf if stop_iteration is None else stop_iteration
        self.stop_time = math.inf if stop_time is None else stop_time
        self.wall_time_limit = math.inf if wall_time_limit is None else wall_time_limit
        self.align_time_step, self.minimum_relative_step, self.verbose = align_time_step, minimum_relative_step, verbose
        self.output_writers = {}
        self.callbacks = {"stop_time_exceeded": Callback(_stop_time_exceeded),
                          "stop_iteration_exceeded": Callback(_stop_iteration_exceeded),
                          "wall_time_limit_exceeded": Callback(_wall_time_limit_exceeded),
                          # default_nan_checker: the first prognostic field (Models.jl:201-213)
                          "nan_checker": Callback(NaNChecker({"u": model.velocities["u"]}), IterationInterval(100))}
        self.running = False
        self.initialized = False
        self.run_wall_time = 0.0

    def add_callback(self, name, func, iterations=1):
        self.callbacks[name] = Callback(func, IterationInterval(iterations))

    def _activities(self):
        return list(self.callbacks.values()) + list(self.output_writers.values())

    def aligned_time_step(self, dt):
        clock = self.model.clock
        a = dt
        for act in self._activities():
            a = act.schedule.aligned_time_step(clock, a)
        a = min(a, self.stop_time - clock.time)
        return dt if a <= 0 else a

    def initialize(self):
        model = self.model
        if model.clock.iteration == 0:
            for act in self._activities():
                act.schedule.initialize(model)
        model.update_state()
        if model.clock.iteration == 0:
            for cb in self.callbacks.values():
                cb(self)
            for w in self.output_writers.values():
                w.write(self)
        self.initialized = True

    def time_step(self):
        t0 = _time.time()
        if not self.initialized:
            self.initialize()
        dt = self.aligned_time_step(self.Δt) if self.align_time_step else self.Δt
        if dt < self.minimum_relative_step * self.Δt:
            t, it, st, ldt, lsdt = self.model.clock._get()
            from . import _abi
            _abi.check(self.model._lib.om_set_clock(self.model._h, t + dt, it, st, ldt, lsdt), self.model._h)
        else:
            self.model.time_step(dt, 1)
        for cb in self.callbacks.values():
            if cb.schedule(self.model):
                cb(self)
        for w in self.output_writers.values():
            if w.schedule(self.model):
                w.write(self)
        self.run_wall_time += _time.time() - t0


def _nan_detected(sim):
    cb = sim.callbacks.get("nan_checker")
    return cb is not None and getattr(cb.func, "nan_detected", False)


def set_simulation_(sim, checkpoint=None, iteration=None):
    """set!(simulation; checkpoint, iteration): restore the prognostic state from a checkpoint
    (run.jl:68-101).  checkpoint: a path, "latest" / "recent_time_stamp" or "highest_iteration"."""
    if checkpoint is None and iteration is None:
        return
    cps = [w for w in sim.output_writers.values() if isinstance(w, Checkpointer)]
    if iteration is not None:
        path = cps[0].path(iteration)
    elif checkpoint in ("latest", "recent_time_stamp"):
        path = max((p for _, p in cps[0].paths()), key=os.path.getmtime)
    elif checkpoint == "highest_iteration":
        path = cps[0].paths()[-1][1]
    else:
        path = checkpoint
    restore_prognostic_state_(sim.model, load_checkpoint_state(path))


def run_(sim, pickup=False, checkpoint_at_end=False):
    """run!(simulation; pickup, checkpoint_at_end) (run.jl:151-187)."""
    if pickup is not False:
        if pickup is True:
            set_simulation_(sim, checkpoint="recent_time_stamp")
        elif isinstance(pickup, int):
            set_simulation_(sim, iteration=pickup)
        else:
            set_simulation_(sim, checkpoint=pickup)
    sim.initialized = False
    sim.running = True
    sim.run_wall_time = 0.0
    cb = sim.callbacks.get("nan_checker")
    if cb is not None and hasattr(cb.func, "nan_detected"):
        cb.func.nan_detected = False
    while sim.running:
        sim.time_step()
    if checkpoint_at_end and not _nan_detected(sim):
        for w in sim.output_writers.values():
            if isinstance(w, Checkpointer):
                w.write(sim)

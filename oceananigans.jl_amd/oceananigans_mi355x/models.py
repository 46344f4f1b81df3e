"""NonhydrostaticModel on the MI355X (src/Models/NonhydrostaticModels/nonhydrostatic_model.jl:124-313).

``build_config`` turns the reference-style description into the C ABI's ``om_config_t``;
``NonhydrostaticModel`` owns the library handle.  ``set_`` follows ``set!(model; kwargs...)``
(set_nonhydrostatic_model.jl:39-74) and ``time_step_`` follows ``time_step!`` (runge_kutta_3.jl:103-168).
"""
import ctypes
import inspect

import numpy as np

from . import _abi
from .grids import Bounded, Flat, ImmersedBoundaryGrid, Periodic, RectilinearGrid
from .physics import (AnisotropicMinimumDissipation, BuoyancyTracer, Centered, ConjugateGradientPoissonSolver,
                      FFTBasedPoissonSolver, FieldBoundaryConditions, FourierTridiagonalPoissonSolver, FPlane,
                      ScalarDiffusivity, SeawaterBuoyancy, SmagorinskyLilly, WENO)

VELOCITY_NAMES = ("u", "v", "w")


def build_config(grid, advection=None, closure=None, buoyancy=None, coriolis=None, tracers=(),
                 boundary_conditions=None, pressure_solver=None, hydrostatic_pressure_anomaly="default"):
    """Return (om_config_t, keepalive) for a model description.  keepalive holds arrays the
    struct points to (z faces)."""
    if not isinstance(grid, (RectilinearGrid, ImmersedBoundaryGrid)):
        raise TypeError("grid must be a RectilinearGrid or an ImmersedBoundaryGrid")
    if isinstance(tracers, str):
        tracers = (tracers,)
    tracers = tuple(tracers)
    if len(tracers) > _abi.OM_MAX_TRACERS:
        raise ValueError("too many tracers")
    c = _abi.OmConfig()
    c.precision = 4 if grid.float_type == np.float32 else 8
    for d in range(3):
        c.N[d] = grid.N[d]
        c.H[d] = grid.H[d]
        c.topology[d] = grid.topology[d].code
        c.origin[d] = grid.origin[d]
        c.extent[d] = grid.extent[d]
    keep = []
    c.bottom_height = None
    if isinstance(grid, ImmersedBoundaryGrid):
        if advection is not None and not (isinstance(advection, WENO) and advection.order == 5):
            raise ValueError("an immersed bottom needs WENO(order=5) advection (or none)")
        hb = np.ascontiguousarray(grid.bottom_heights().reshape(-1, order="F"), dtype=np.float64)
        keep.append(hb)
        c.bottom_height = hb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if grid.z_faces is not None:
        zf = np.ascontiguousarray(grid.z_faces, dtype=np.float64)
        keep.append(zf)
        c.z_faces = _abi.dptr(zf)
    # advection (reference default: Centered(order=2))
    if advection is None:
        advection = Centered()
    if advection is False:
        c.advection = _abi.OM_ADV_NONE
    elif isinstance(advection, WENO):
        c.advection = _abi.OM_ADV_WENO5
    elif isinstance(advection, Centered):
        c.advection = _abi.OM_ADV_CENTERED2
    else:
        raise ValueError(f"unsupported advection {advection!r}")
    # closure
    if closure is None:
        c.closure = _abi.OM_CLOSURE_NONE
    elif isinstance(closure, AnisotropicMinimumDissipation):
        c.closure = _abi.OM_CLOSURE_AMD
        c.closure_params[0] = closure.Cnu
        c.closure_params[1] = closure.Ckappa
    elif isinstance(closure, SmagorinskyLilly):
        c.closure = _abi.OM_CLOSURE_SMAGORINSKY_LILLY
        c.closure_params[0] = closure.C
        c.closure_params[1] = closure.Cb
        c.closure_params[2] = closure.Pr
    elif isinstance(closure, ScalarDiffusivity):
        c.closure = _abi.OM_CLOSURE_SCALAR_DIFFUSIVITY
        c.closure_params[0] = closure.nu
        c.closure_params[1] = closure.kappa
    else:
        raise ValueError(f"unsupported closure {closure!r}")
    # buoyancy
    c.gravitational_acceleration = 9.80665
    if buoyancy is None:
        c.buoyancy = _abi.OM_BUOYANCY_NONE
    elif isinstance(buoyancy, BuoyancyTracer):
        if "b" not in tracers:
            raise ValueError("BuoyancyTracer requires a tracer named b")
        c.buoyancy = _abi.OM_BUOYANCY_TRACER
        c.buoyancy_tracer = tracers.index("b")
    elif isinstance(buoyancy, SeawaterBuoyancy):
        if "T" not in tracers or "S" not in tracers:
            raise ValueError("SeawaterBuoyancy requires tracers T and S")
        c.buoyancy = _abi.OM_BUOYANCY_LINEAR_EOS
        c.T_tracer = tracers.index("T")
        c.S_tracer = tracers.index("S")
        c.gravitational_acceleration = buoyancy.gravitational_acceleration
        c.thermal_expansion = buoyancy.equation_of_state.thermal_expansion
        c.haline_contraction = buoyancy.equation_of_state.haline_contraction
    else:
        raise ValueError(f"unsupported buoyancy {buoyancy!r}")
    # Coriolis
    if coriolis is None:
        c.coriolis = 0
    elif isinstance(coriolis, FPlane):
        c.coriolis = 1
        c.f = coriolis.f
    else:
        raise ValueError(f"unsupported coriolis {coriolis!r}")
    c.n_tracers = len(tracers)
    # boundary conditions (z only)
    bcs = boundary_conditions or {}
    names = VELOCITY_NAMES + tracers
    for name, fbc in bcs.items():
        if name not in names:
            raise ValueError(f"boundary conditions given for unknown field {name}")
        if not isinstance(fbc, FieldBoundaryConditions):
            raise TypeError("boundary conditions must be FieldBoundaryConditions")
        f = names.index(name)
        if name == "w" and (fbc.top is not None or fbc.bottom is not None):
            raise ValueError("w keeps its impenetrable top and bottom")
        if grid.topology[2] is not Bounded and (fbc.top is not None or fbc.bottom is not None):
            raise ValueError("top/bottom boundary conditions need a Bounded z")
        if fbc.bottom is not None:
            c.bc_bottom[f] = fbc.bottom.kind
            c.bc_bottom_value[f] = fbc.bottom.condition
        if fbc.top is not None:
            c.bc_top[f] = fbc.top.kind
            c.bc_top_value[f] = fbc.top.condition
    # pressure solver (NonhydrostaticModels.jl:40-56: FFT for uniform grids, Fourier–tridiagonal for stretched z)
    c.cg_reltol, c.cg_abstol, c.cg_maxiter, c.cg_preconditioner = -1.0, -1.0, 0, _abi.OM_PRECOND_DEFAULT
    if pressure_solver is None:
        c.pressure_solver = _abi.OM_SOLVER_FOURIER_TRIDIAGONAL if grid.stretched else _abi.OM_SOLVER_FFT
    elif isinstance(pressure_solver, FFTBasedPoissonSolver) or pressure_solver in ("FFTBasedPoissonSolver", _abi.OM_SOLVER_FFT):
        c.pressure_solver = _abi.OM_SOLVER_FFT
    elif isinstance(pressure_solver, FourierTridiagonalPoissonSolver) or \
            pressure_solver in ("FourierTridiagonalPoissonSolver", _abi.OM_SOLVER_FOURIER_TRIDIAGONAL):
        c.pressure_solver = _abi.OM_SOLVER_FOURIER_TRIDIAGONAL
    elif isinstance(pressure_solver, ConjugateGradientPoissonSolver):
        if any(t is Flat for t in grid.topology):
            raise ValueError("ConjugateGradientPoissonSolver needs a three-dimensional grid")
        c.pressure_solver = _abi.OM_SOLVER_CONJUGATE_GRADIENT
        c.cg_preconditioner = pressure_solver.preconditioner_kind()
        if pressure_solver.reltol is not None:
            c.cg_reltol = float(pressure_solver.reltol)
        if pressure_solver.abstol is not None:
            c.cg_abstol = float(pressure_solver.abstol)
        if pressure_solver.maxiter is not None:
            c.cg_maxiter = int(pressure_solver.maxiter)
    else:
        raise ValueError(f"unsupported pressure solver {pressure_solver!r}")
    c.hydrostatic_pressure = {"default": -1, None: 0, False: 0, True: 1}[hydrostatic_pressure_anomaly]
    return c, keep


def interior_shape(grid, field_id):
    nz = grid.Nz
    if field_id % 100 == 2 and field_id < 300 and grid.topology[2] is Bounded:
        nz += 1
    return (grid.Nx, grid.Ny, nz)


def field_nodes(grid, name):
    """Node coordinates (x, y, z) of a prognostic field (u: fcc, v: cfc, w: ccf, tracers: ccc)."""
    x = grid.xnodes("f" if name == "u" else "c")
    y = grid.ynodes("f" if name == "v" else "c")
    z = grid.znodes("f" if name == "w" else "c")
    return x, y, z


class Clock:
    def __init__(self, model):
        self._m = model

    def _get(self):
        t, it, st, ldt, lsdt = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        _abi.check(self._m._lib.om_get_clock(self._m._h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(st),
                                             ctypes.byref(ldt), ctypes.byref(lsdt)), self._m._h)
        return t.value, it.value, st.value, ldt.value, lsdt.value

    time = property(lambda s: s._get()[0])
    iteration = property(lambda s: s._get()[1])
    stage = property(lambda s: s._get()[2])
    last_Δt = property(lambda s: s._get()[3])


class Field:
    """A device field of the model; ``interior()`` copies it to the host (FP64)."""

    def __init__(self, model, field_id, name):
        self.model, self.id, self.name = model, field_id, name

    def interior(self):
        shape = interior_shape(self.model.grid, self.id)
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _abi.check(self.model._lib.om_get_field(self.model._h, self.id, _abi.dptr(out), 0), self.model._h)
        return out.reshape(shape, order="F")

    def parent(self):
        info = self.info()
        shape = tuple(info.parent_size)
        out = np.empty(int(np.prod(shape)), dtype=np.float64)
        _abi.check(self.model._lib.om_get_field(self.model._h, self.id, _abi.dptr(out), 1), self.model._h)
        return out.reshape(shape, order="F")

    def maximum(self):
        """maximum(field.data.parent) (halos included), reduced on the device."""
        return self.model.field_statistics(self.id, True)[0]

    def minimum(self):
        return self.model.field_statistics(self.id, True)[1]

    def hasnan(self):
        """any(isnan, parent(field)) (src/Diagnostics/nan_checker.jl:37)."""
        return self.model.field_statistics(self.id, True)[2] > 0

    def set_parent(self, values):
        """copyto!(parent(field), values): the whole padded array, halos included."""
        buf = _abi.as_fortran_f64(np.asarray(values, dtype=np.float64))
        _abi.check(self.model._lib.om_set_field(self.model._h, self.id, _abi.dptr(buf), 1), self.model._h)

    def info(self):
        info = _abi.OmFieldInfo()
        _abi.check(self.model._lib.om_field_info(self.model._h, self.id, ctypes.byref(info)), self.model._h)
        return info

    def __repr__(self):
        return f"Field {self.name} on {self.model.grid!r}"


class _NamedFields(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e


class NonhydrostaticModel:
    def __init__(self, grid, *, advection=None, closure=None, buoyancy=None, coriolis=None, tracers=(),
                 boundary_conditions=None, pressure_solver=None, hydrostatic_pressure_anomaly="default",
                 forcing=None, stokes_drift=None, background_fields=None, particles=None, biogeochemistry=None,
                 timestepper="RungeKutta3"):
        for name, val in (("forcing", forcing), ("stokes_drift", stokes_drift), ("background_fields", background_fields),
                          ("particles", particles), ("biogeochemistry", biogeochemistry)):
            if val:
                raise ValueError(f"{name} is not supported on the MI355X path")
        if timestepper != "RungeKutta3":
            raise ValueError("the MI355X path implements timestepper = :RungeKutta3")
        if isinstance(tracers, str):
            tracers = (tracers,)
        self.grid = grid
        self.tracer_names = tuple(tracers)
        self.advection, self.closure, self.buoyancy, self.coriolis = advection, closure, buoyancy, coriolis
        self.config, self._keep = build_config(grid, advection, closure, buoyancy, coriolis, self.tracer_names,
                                               boundary_conditions, pressure_solver, hydrostatic_pressure_anomaly)
        self._lib = _abi.load_library()
        h = ctypes.c_void_p()
        rc = self._lib.om_create(ctypes.byref(self.config), ctypes.byref(h))
        if rc != 0:
            raise _abi.OmError(f"{_abi.ERRORS.get(rc, rc)}: {self._lib.om_last_error(None).decode()}")
        self._h = h
        names = VELOCITY_NAMES + self.tracer_names
        self.velocities = _NamedFields((n, Field(self, i, n)) for i, n in enumerate(VELOCITY_NAMES))
        self.tracers = _NamedFields((n, Field(self, 3 + i, n)) for i, n in enumerate(self.tracer_names))
        self.pressures = _NamedFields(pNHS=Field(self, _abi.OM_FIELD_PNHS, "pNHS"),
                                      pHY=Field(self, _abi.OM_FIELD_PHY, "pHY′"))
        self.closure_fields = _NamedFields([("nu", Field(self, _abi.OM_FIELD_NU, "νₑ"))] +
                                           [("kappa_" + n, Field(self, _abi.OM_FIELD_KAPPA0 + i, "κₑ " + n))
                                            for i, n in enumerate(self.tracer_names)])
        self.Gn = _NamedFields((n, Field(self, _abi.OM_FIELD_GN + i, "G" + n)) for i, n in enumerate(names))
        self.Gm = _NamedFields((n, Field(self, _abi.OM_FIELD_GM + i, "G⁻" + n)) for i, n in enumerate(names))
        self.clock = Clock(self)
        pc = getattr(pressure_solver, "preconditioner", None)
        if getattr(pc, "weights", None) is not None:
            self.learned_set_weights(pc.weights)
        # the constructor ends with update_state!(model) (nonhydrostatic_model.jl:309)
        self.update_state()

    # ---------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, "_h", None):
            self._lib.om_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- reference operations
    def field_id(self, name):
        names = VELOCITY_NAMES + self.tracer_names
        if name not in names:
            raise ValueError(f"name {name} not found in model.velocities or model.tracers.")
        return names.index(name)

    def set_interior(self, name, values):
        fid = self.field_id(name) if isinstance(name, str) else name
        shape = interior_shape(self.grid, fid)
        a = np.asarray(values, dtype=np.float64)
        if a.shape != shape:
            if fid == 2 and a.shape == (shape[0], shape[1], shape[2] - 1):
                a = np.concatenate([a, np.zeros((shape[0], shape[1], 1))], axis=2)
            else:
                a = np.broadcast_to(a, shape)
        buf = _abi.as_fortran_f64(a)
        _abi.check(self._lib.om_set_field(self._h, fid, _abi.dptr(buf), 0), self._h)

    def fill_halo_regions(self, names=None, fill_normal_flow_bcs=True):
        names = names or (VELOCITY_NAMES + self.tracer_names)
        mask = 0
        for n in names:
            mask |= 1 << self.field_id(n)
        _abi.check(self._lib.om_fill_halos(self._h, mask, 1 if fill_normal_flow_bcs else 0), self._h)

    def update_state(self):
        _abi.check(self._lib.om_update_state(self._h), self._h)

    def compute_tendencies(self):
        """compute_tendencies!(model) on the current state, closure fields and pHY′ (no halo fill)."""
        _abi.check(self._lib.om_compute_tendencies(self._h), self._h)

    def compute_pressure_correction(self, dt):
        _abi.check(self._lib.om_pressure_correction(self._h, float(dt)), self._h)

    def time_step(self, dt, nsteps=1):
        if dt == 0:
            import warnings
            warnings.warn("Δt == 0 may cause model blowup!")
        _abi.check(self._lib.om_time_step(self._h, float(dt), int(nsteps)), self._h)

    def autotune(self):
        """Pick the z-chunk length of the tiled kernels by timing them on the current state."""
        kc = ctypes.c_int32()
        _abi.check(self._lib.om_autotune(self._h, ctypes.byref(kc)), self._h)
        return kc.value

    def last_solve_info(self):
        """(iterations, ‖V⁻¹r‖ the iteration stopped on, ‖V⁻¹(b − A x)‖ recomputed in FP64) of the last
        conjugate-gradient pressure solve."""
        it, r, t = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        _abi.check(self._lib.om_last_solve_info(self._h, ctypes.byref(it), ctypes.byref(r), ctypes.byref(t)), self._h)
        return it.value, r.value, t.value

    def set_tuning(self, tendency_kernel=None, kchunk=None, learned_iterations=None, poisson_refinement=None,
                   step_graphs=None, conv_tiled=None, wgrad_mfma=None):
        """Launch-shape choices: tendency_kernel 'tiled' (default where supported) or 'generic';
        kchunk = z levels per block of the tiled kernel; step_graphs: time_step replays a captured
        two-step hipGraph (default) or issues every launch; conv_tiled: the learned preconditioner's
        convolutions with LDS-staged input tiles (default) or not.  Results do not depend on them.
        poisson_refinement: one FP64 iterative-refinement step after each FFT / Fourier–tridiagonal
        Poisson solve (off by default)."""
        if wgrad_mfma is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_WGRAD_MFMA, 1 if wgrad_mfma else 0), self._h)
        if conv_tiled is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_CONV_TILED, 1 if conv_tiled else 0), self._h)
        if step_graphs is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_STEP_GRAPHS, 1 if step_graphs else 0), self._h)
        if poisson_refinement is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_POISSON_REFINEMENT, 1 if poisson_refinement else 0), self._h)
        if tendency_kernel is not None:
            v = {"tiled": 0, "generic": 1}[tendency_kernel]
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_TENDENCY_KERNEL, v), self._h)
        if kchunk is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_KCHUNK, int(kchunk)), self._h)
        if learned_iterations is not None:
            _abi.check(self._lib.om_set_tuning(self._h, _abi.OM_TUNE_LEARNED_ITERATIONS, int(learned_iterations)), self._h)

    def get_tuning(self):
        v = ctypes.c_int32()
        _abi.check(self._lib.om_get_tuning(self._h, _abi.OM_TUNE_TENDENCY_KERNEL, ctypes.byref(v)), self._h)
        kernel = ("tiled", "generic")[v.value]
        _abi.check(self._lib.om_get_tuning(self._h, _abi.OM_TUNE_KCHUNK, ctypes.byref(v)), self._h)
        kchunk = v.value
        _abi.check(self._lib.om_get_tuning(self._h, _abi.OM_TUNE_STEP_GRAPHS, ctypes.byref(v)), self._h)
        return {"tendency_kernel": kernel, "kchunk": kchunk, "step_graphs": bool(v.value)}

    def poisson_refinement(self):
        v = ctypes.c_int32()
        _abi.check(self._lib.om_get_tuning(self._h, _abi.OM_TUNE_POISSON_REFINEMENT, ctypes.byref(v)), self._h)
        return bool(v.value)

    def set_phase_timing(self, enable=True):
        _abi.check(self._lib.om_set_phase_timing(self._h, 1 if enable else 0), self._h)

    def last_step_ms(self):
        t = ctypes.c_double()
        _abi.check(self._lib.om_last_step_timing(self._h, ctypes.byref(t), None, None, 0), self._h)
        return t.value

    def phase_times(self):
        """{phase: (total ms, launches)} of the last time_step call with phase timing enabled."""
        n = _abi.OM_NPHASES
        t = ctypes.c_double()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int32 * n)()
        _abi.check(self._lib.om_last_step_timing(self._h, ctypes.byref(t), ms, cnt, n), self._h)
        return {name: (ms[p], cnt[p]) for p, name in enumerate(_abi.PHASES)}

    def cell_advection_timescale(self):
        """cell_advection_timescale(model) (NonhydrostaticModels.jl:93-97): min over cells of
        1 / (|u|/Δx + |v|/Δy + |w|/Δz), reduced on the device."""
        t = ctypes.c_double()
        _abi.check(self._lib.om_cell_advection_timescale(self._h, ctypes.byref(t)), self._h)
        return t.value

    def field_statistics_noalloc(self, field_id, with_halos=True):
        """field_statistics through om_field_statistics_noalloc (no allocation in the call)."""
        mx, mn, nn = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _abi.check(self._lib.om_field_statistics_noalloc(self._h, int(field_id), 1 if with_halos else 0, ctypes.byref(mx),
                                                         ctypes.byref(mn), ctypes.byref(nn)), self._h)
        return mx.value, mn.value, nn.value

    def field_statistics(self, field_id, with_halos=True):
        """(maximum, minimum, number of NaNs) of a field's parent data (or interior), on the device."""
        mx, mn, nn = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _abi.check(self._lib.om_field_statistics(self._h, int(field_id), 1 if with_halos else 0, ctypes.byref(mx),
                                                 ctypes.byref(mn), ctypes.byref(nn)), self._h)
        return mx.value, mn.value, nn.value

    # ---------------------------------------------------------------- learned preconditioner
    def learned_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        if w.size != _abi.OM_LEARNED_NWEIGHTS:
            raise ValueError(f"expected {_abi.OM_LEARNED_NWEIGHTS} weights")
        _abi.check(self._lib.om_learned_set_weights(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), self._h)

    def learned_get_weights(self):
        w = np.empty(_abi.OM_LEARNED_NWEIGHTS, dtype=np.float32)
        _abi.check(self._lib.om_learned_get_weights(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), self._h)
        return w.reshape(_abi.OM_LEARNED_LAYERS, 27, 16, 16)

    def learned_log(self, iterations_per_solve=4, capacity=64):
        _abi.check(self._lib.om_learned_log(self._h, int(iterations_per_solve), int(capacity)), self._h)

    def learned_pairs(self):
        n = ctypes.c_int32()
        _abi.check(self._lib.om_learned_pairs(self._h, ctypes.byref(n)), self._h)
        return n.value

    def learned_pairs_logged(self):
        """Pairs logged since learned_log (the ring holds the last `capacity` of them)."""
        n = ctypes.c_int64()
        _abi.check(self._lib.om_learned_pairs_logged(self._h, ctypes.byref(n)), self._h)
        return n.value

    def learned_train(self, steps, lr=1e-3):
        a, b = ctypes.c_double(), ctypes.c_double()
        _abi.check(self._lib.om_learned_train(self._h, int(steps), float(lr), ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, b.value

    def learned_evaluate(self):
        """Mean relative A-energy of the remaining error over the held pairs (1: the FFT
        preconditioner alone)."""
        x = ctypes.c_double()
        _abi.check(self._lib.om_learned_evaluate(self._h, ctypes.byref(x)), self._h)
        return x.value

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def learned_forward(self, input3):
        """Network output (Nx, Ny, Nz) for a (Nx, Ny, Nz, 3) input."""
        g = self.grid
        x = np.ascontiguousarray(np.transpose(np.asarray(input3, dtype=np.float32), (2, 1, 0, 3)))
        out = np.empty((g.Nz, g.Ny, g.Nx), dtype=np.float32)
        _abi.check(self._lib.om_learned_forward(self._h, self._fp(x), self._fp(out)), self._h)
        return np.transpose(out, (2, 1, 0))

    def learned_backward(self, input3, target, mask):
        """(loss, dL/dW (7, 27, 16, 16), dL/d(input) (Nx, Ny, Nz, 3)) of the mean-square loss."""
        g = self.grid
        tr = lambda a: np.ascontiguousarray(np.transpose(np.asarray(a, dtype=np.float32), (2, 1, 0) + tuple(range(3, np.ndim(a)))))
        x, t, m = tr(input3), tr(target), tr(mask)
        dw = np.empty(_abi.OM_LEARNED_NWEIGHTS, dtype=np.float32)
        din = np.empty((g.Nz, g.Ny, g.Nx, 3), dtype=np.float32)
        loss = ctypes.c_double()
        _abi.check(self._lib.om_learned_backward(self._h, self._fp(x), self._fp(t), self._fp(m), ctypes.byref(loss),
                                                 self._fp(dw), self._fp(din)), self._h)
        return loss.value, dw.reshape(_abi.OM_LEARNED_LAYERS, 27, 16, 16), np.transpose(din, (2, 1, 0, 3))

    def learned_energy_loss(self, output, target):
        """(E, dE/dy) of the training loss's energy for a network output y and target t, both
        (Nx, Ny, Nz): d = active·(y − t), E = dᵀ(−V∇²)d with the CG solver's operator."""
        g = self.grid
        tr = lambda a: np.ascontiguousarray(np.transpose(np.asarray(a, dtype=np.float32), (2, 1, 0)))
        y, t = tr(output), tr(target)
        grad = np.empty((g.Nz, g.Ny, g.Nx), dtype=np.float32)
        e = ctypes.c_double()
        _abi.check(self._lib.om_learned_energy_loss(self._h, self._fp(y), self._fp(t), ctypes.byref(e), self._fp(grad)), self._h)
        return e.value, np.transpose(grad, (2, 1, 0))

    def learned_last_solve(self):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _abi.check(self._lib.om_learned_last_solve(self._h, ctypes.byref(a), ctypes.byref(b)), self._h)
        return a.value, bool(b.value)

    def divergence(self):
        n = self.grid.Nx * self.grid.Ny * self.grid.Nz
        out = np.empty(n)
        _abi.check(self._lib.om_divergence(self._h, _abi.dptr(out)), self._h)
        return out.reshape((self.grid.Nx, self.grid.Ny, self.grid.Nz), order="F")

    def poisson_solve(self, rhs):
        rhs = _abi.as_fortran_f64(rhs)
        out = np.empty_like(rhs)
        _abi.check(self._lib.om_poisson_solve(self._h, _abi.dptr(rhs), _abi.dptr(out)), self._h)
        return out.reshape((self.grid.Nx, self.grid.Ny, self.grid.Nz), order="F")

    def __repr__(self):
        return (f"NonhydrostaticModel{{MI355X, RectilinearGrid}}(time = {self.clock.time}, iteration = {self.clock.iteration})\n"
                f"├── grid: {self.grid!r}\n├── timestepper: RungeKutta3TimeStepper\n"
                f"├── advection scheme: {self.advection!r}\n├── tracers: {self.tracer_names}\n"
                f"├── closure: {self.closure!r}\n├── buoyancy: {type(self.buoyancy).__name__ if self.buoyancy else 'Nothing'}\n"
                f"└── coriolis: {type(self.coriolis).__name__ if self.coriolis else 'Nothing'}")


def _evaluate(grid, name, value):
    """Field values for set!: arrays are copied; functions are evaluated at the field's nodes
    (src/Fields/set!.jl:81-125), f(x, y, z) in 3D and f(x, y) when z is Flat."""
    if not callable(value):
        return value
    x, y, z = field_nodes(grid, name)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    if grid.topology[2] is Flat:
        nargs = len(inspect.signature(value).parameters)
        out = value(X, Y) if nargs == 2 else value(X, Y, Z)
    else:
        out = value(X, Y, Z)
    return np.broadcast_to(np.asarray(out, dtype=np.float64), X.shape)


def set_(model, enforce_incompressibility=True, **kwargs):
    """set!(model; kwargs...) — set_nonhydrostatic_model.jl:39-74."""
    for name, value in kwargs.items():
        model.set_interior(name, _evaluate(model.grid, name, value))
        model.fill_halo_regions([name], fill_normal_flow_bcs=True)
    model.update_state()
    if enforce_incompressibility:
        model.compute_pressure_correction(1.0)
        model.update_state()


def time_step_(model, dt):
    """time_step!(model, Δt) — one RK3 step (runge_kutta_3.jl:103-168)."""
    model.time_step(dt, 1)


class Ensemble:
    """Independent models advanced together (no NonhydrostaticModel counterpart in the reference,
    where members are stepped one after another).  Members keep their own sizes.

    Forms (``mode``): "batched" merges every member's launch of each kernel of the step into one
    launch over a table of the members (one grid, each member a contiguous block range) and replays
    the whole two-step sequence as one hipGraph; members must share one launch sequence (same
    precision, closure, buoyancy, tracers, FFT solver).  "member_graphs" replays each member's own
    two-step hipGraph on its own stream, members issued in rounds so that their streams run
    concurrently; "direct" issues the same launches member by member."""

    MODES = {"direct": _abi.OM_ENSEMBLE_DIRECT, "member_graphs": _abi.OM_ENSEMBLE_MEMBER_GRAPHS, "batched": _abi.OM_ENSEMBLE_BATCHED}

    def __init__(self, models):
        self.models = list(models)
        self._lib = _abi.load_library()
        arr = (ctypes.c_void_p * len(self.models))(*[m._h for m in self.models])
        h = ctypes.c_void_p()
        rc = self._lib.om_ensemble_create(arr, len(self.models), ctypes.byref(h))
        if rc != 0:
            raise _abi.OmError(f"{_abi.ERRORS.get(rc, rc)}: {self._lib.om_ensemble_last_error(None).decode()}")
        self._h = h

    def set_batch_groups(self, groups):
        """The batched form's member groups, each a member-table graph on its own stream."""
        rc = self._lib.om_ensemble_set_batch_groups(self._h, int(groups))
        if rc != 0:
            raise _abi.OmError(f"{_abi.ERRORS.get(rc, rc)}: invalid group count {groups}")

    def time_step(self, dt, nsteps=1, graphs=True, mode=None):
        if mode is None:
            mode = "member_graphs" if graphs else "direct"
        rc = self._lib.om_ensemble_time_step(self._h, float(dt), int(nsteps), self.MODES[mode])
        if rc != 0:
            raise _abi.OmError(f"{_abi.ERRORS.get(rc, rc)}: {self._lib.om_ensemble_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._lib.om_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes mirror of include/oceananigans_mi355x.h and the loader of the HIP library.

The library is loaded from the package directory (built in-tree by ``build.py``).  There is no
CPU fallback: if the library is missing or no GPU is visible, the calls fail loudly.
"""
import ctypes
import os

import numpy as np

OM_PERIODIC, OM_BOUNDED, OM_FLAT = 0, 1, 2
OM_ADV_NONE, OM_ADV_CENTERED2, OM_ADV_WENO5 = 0, 1, 2
OM_CLOSURE_NONE, OM_CLOSURE_AMD, OM_CLOSURE_SMAGORINSKY_LILLY, OM_CLOSURE_SCALAR_DIFFUSIVITY = 0, 1, 2, 3
OM_BUOYANCY_NONE, OM_BUOYANCY_TRACER, OM_BUOYANCY_LINEAR_EOS = 0, 1, 2
OM_BC_DEFAULT, OM_BC_FLUX, OM_BC_VALUE, OM_BC_GRADIENT = 0, 1, 2, 3
OM_MAX_TRACERS = 8
OM_MAX_FIELDS = 3 + OM_MAX_TRACERS

OM_FIELD_U, OM_FIELD_V, OM_FIELD_W, OM_FIELD_TRACER0 = 0, 1, 2, 3
OM_FIELD_GN, OM_FIELD_GM = 100, 200
OM_FIELD_PNHS, OM_FIELD_PHY, OM_FIELD_NU, OM_FIELD_KAPPA0 = 300, 301, 302, 303
OM_SOLVER_FFT, OM_SOLVER_FOURIER_TRIDIAGONAL, OM_SOLVER_CONJUGATE_GRADIENT = 0, 1, 2
OM_PRECOND_DEFAULT, OM_PRECOND_NONE, OM_PRECOND_DIAGONALLY_DOMINANT, OM_PRECOND_FFT = 0, 1, 2, 3
OM_PRECOND_COLUMNWISE_TRIDIAGONAL = 4
OM_PRECOND_LEARNED = 5
OM_LEARNED_LAYERS = 7
OM_LEARNED_NWEIGHTS = OM_LEARNED_LAYERS * 27 * 16 * 16
OM_TUNE_TENDENCY_KERNEL = 0
OM_TUNE_KCHUNK = 1
OM_TUNE_LEARNED_ITERATIONS = 2
OM_TUNE_POISSON_REFINEMENT = 3
OM_TUNE_STEP_GRAPHS = 4
OM_TUNE_CONV_TILED = 5
OM_TUNE_WGRAD_MFMA = 6
OM_ENSEMBLE_DIRECT, OM_ENSEMBLE_MEMBER_GRAPHS, OM_ENSEMBLE_BATCHED = 0, 1, 2
PHASES = ("rk_update", "poisson", "correction", "halos", "closure", "hydrostatic", "tendencies")
OM_NPHASES = len(PHASES)

OM_EINVAL, OM_EHIP, OM_ENOMEM, OM_ESTATE, OM_ENOTCONV = -1, -2, -3, -4, -5
ERRORS = {-1: "OM_EINVAL", -2: "OM_EHIP", -3: "OM_ENOMEM", -4: "OM_ESTATE", -5: "OM_ENOTCONV"}


class OmConfig(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("N", ctypes.c_int32 * 3),
        ("H", ctypes.c_int32 * 3),
        ("topology", ctypes.c_int32 * 3),
        ("origin", ctypes.c_double * 3),
        ("extent", ctypes.c_double * 3),
        ("z_faces", ctypes.POINTER(ctypes.c_double)),
        ("advection", ctypes.c_int32),
        ("closure", ctypes.c_int32),
        ("closure_params", ctypes.c_double * 4),
        ("buoyancy", ctypes.c_int32),
        ("buoyancy_tracer", ctypes.c_int32),
        ("T_tracer", ctypes.c_int32),
        ("S_tracer", ctypes.c_int32),
        ("gravitational_acceleration", ctypes.c_double),
        ("thermal_expansion", ctypes.c_double),
        ("haline_contraction", ctypes.c_double),
        ("coriolis", ctypes.c_int32),
        ("f", ctypes.c_double),
        ("n_tracers", ctypes.c_int32),
        ("bc_bottom", ctypes.c_int32 * OM_MAX_FIELDS),
        ("bc_top", ctypes.c_int32 * OM_MAX_FIELDS),
        ("bc_bottom_value", ctypes.c_double * OM_MAX_FIELDS),
        ("bc_top_value", ctypes.c_double * OM_MAX_FIELDS),
        ("pressure_solver", ctypes.c_int32),
        ("hydrostatic_pressure", ctypes.c_int32),
        ("cg_reltol", ctypes.c_double),
        ("cg_abstol", ctypes.c_double),
        ("cg_maxiter", ctypes.c_int32),
        ("cg_preconditioner", ctypes.c_int32),
        ("bottom_height", ctypes.POINTER(ctypes.c_double)),
        ("reserved", ctypes.c_int32 * 8),
    ]


class OmFieldInfo(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("element_size", ctypes.c_int32),
        ("parent_size", ctypes.c_int32 * 3),
        ("offset", ctypes.c_int32 * 3),
        ("stride", ctypes.c_int64 * 3),
    ]


LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "liboceananigans_mi355x.so")

# every entry point declared in include/oceananigans_mi355x.h, with its ctypes signature
SIGNATURES = {
    "om_abi_version": (ctypes.c_int, []),
    "om_create": (ctypes.c_int, [ctypes.POINTER(OmConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "om_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "om_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "om_field_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OmFieldInfo)]),
    "om_set_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    "om_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    "om_fill_halos": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32]),
    "om_update_state": (ctypes.c_int, [ctypes.c_void_p]),
    "om_compute_tendencies": (ctypes.c_int, [ctypes.c_void_p]),
    "om_rk3_substep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32]),
    "om_cache_previous_tendencies": (ctypes.c_int, [ctypes.c_void_p]),
    "om_pressure_correction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "om_time_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    "om_get_clock": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    "om_set_clock": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_double]),
    "om_poisson_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "om_divergence": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "om_set_phase_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "om_last_step_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "om_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "om_last_solve_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]),
    "om_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "om_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "om_ensemble_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "om_ensemble_time_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]),
    "om_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "om_ensemble_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "om_ensemble_set_batch_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "om_cell_advection_timescale": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "om_learned_set_weights": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "om_learned_get_weights": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "om_learned_log": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "om_learned_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "om_learned_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]),
    "om_learned_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "om_learned_evaluate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "om_learned_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "om_learned_last_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "om_learned_energy_loss": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "om_learned_pairs_logged": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "om_field_statistics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "om_field_statistics_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                   ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library (no fallback: raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class OmError(RuntimeError):
    pass


class SolverNotConvergedWarning(RuntimeWarning):
    """A conjugate-gradient pressure solve stopped at maxiter above its tolerance; the call's work
    completed (the reference's solver stops there silently, conjugate_gradient_solver.jl:279-284)."""


def check(rc, handle=None):
    if rc != 0:
        lib = load_library()
        msg = lib.om_last_error(handle)
        msg = msg.decode() if msg else ""
        if rc == OM_ENOTCONV:
            import warnings
            warnings.warn(msg, SolverNotConvergedWarning, stacklevel=3)
            return
        raise OmError(f"{ERRORS.get(rc, rc)}: {msg}")


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def as_fortran_f64(a):
    """Flatten an (Nx, Ny, Nz) array indexed [i, j, k] into x-fastest FP64."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))

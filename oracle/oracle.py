"""ctypes harness of the CPU oracle (om_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg.  The product path never imports it.

The oracle takes the same ``om_config_t`` as the HIP library (built by
``oceananigans_mi355x.build_config``) and restates the reference algorithm on the CPU in FP64
(or FP32 for rounding studies), one reference kernel at a time.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(plant=0):
    """The oracle library; plant > 0 loads the build with that planted mistake (om_oracle.cpp
    ORC_PLANT), which only the tests that show a pin fails on it use."""
    if plant in _libs:
        return _libs[plant]
    # ORACLE_LIBRARY names another build of the oracle (the sanitizer build in its own test)
    default = os.environ.get("ORACLE_LIBRARY", LIB)
    path = default if plant == 0 else os.path.join(HERE, f"liboracle_plant{plant}.so")
    if plant == 0 and path == LIB and not os.path.exists(LIB):
        build()
    if plant != 0:
        subprocess.run(["make", "-s", "-C", HERE, os.path.basename(path)], check=True)
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    pd = P(d)
    sig = {
        "orc_create": (vp, [vp]),
        "orc_destroy": (None, [vp]),
        "orc_set_field": (i, [vp, i, pd]),
        "orc_get_field": (i, [vp, i, pd]),
        "orc_get_field_range": (i, [vp, i, P(i), P(i), pd]),
        "orc_set_field_range": (i, [vp, i, P(i), P(i), pd]),
        "orc_fill_halos": (None, [vp, i]),
        "orc_update_state": (None, [vp]),
        "orc_compute_closure": (None, [vp]),
        "orc_compute_hydrostatic": (None, [vp]),
        "orc_compute_tendencies": (None, [vp]),
        "orc_rk3_substep": (None, [vp, d, d, d, i]),
        "orc_cache_previous_tendencies": (None, [vp]),
        "orc_pressure_correction": (None, [vp, d]),
        "orc_time_step": (None, [vp, d, i]),
        "orc_set_finalize": (None, [vp, i]),
        "orc_get_clock": (None, [vp, pd, P(ctypes.c_int64), P(i)]),
        "orc_set_clock": (None, [vp, d, ctypes.c_int64, i, d, d]),
        "orc_set_threads": (None, [i]),
        "orc_max_threads": (i, []),
        "orc_divergence": (None, [vp, pd]),
        "orc_poisson_solve": (None, [vp, pd, pd]),
        "orc_last_solve_info": (None, [vp, P(i), pd, pd]),
        "orc_cg_precondition": (None, [vp, i, pd, pd]),
        "orc_cell_advection_timescale": (d, [vp]),
        "orc_grid_z": (None, [vp, pd, pd, pd, pd]),
        "orc_weno5": (d, [pd, i, i]),
        "orc_weno3": (d, [pd, i, i]),
        "orc_centered4": (d, [pd, i]),
        "orc_weno5_smoothness": (None, [pd, i, i, pd, pd]),
        "orc_coefficients": (None, [i, pd, pd, pd, pd]),
        "orc_thomas": (None, [i, i, pd, pd, pd, pd, pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _libs[plant] = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


class OracleModel:
    """CPU restatement of NonhydrostaticModel on the same om_config_t."""

    def __init__(self, config, grid, plant=0):
        self.L = lib(plant)
        self.config = config
        self.grid = grid
        self.h = self.L.orc_create(ctypes.byref(config))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self.nt = config.n_tracers

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, fid):
        g = self.grid
        nz = g.Nz
        if fid < 300 and fid % 100 == 2 and g.topology[2].code == 1:
            nz += 1
        return (g.Nx, g.Ny, nz)

    def set(self, fid, values):
        shape = self.shape(fid)
        a = np.asarray(values, dtype=np.float64)
        if a.shape != shape:
            if fid == 2 and a.shape == (shape[0], shape[1], shape[2] - 1):
                a = np.concatenate([a, np.zeros((shape[0], shape[1], 1))], axis=2)
            else:
                a = np.broadcast_to(a, shape)
        buf = _f64(a)
        if self.L.orc_set_field(self.h, fid, _p(buf)) != 0:
            raise ValueError(f"unknown field {fid}")

    def get(self, fid):
        shape = self.shape(fid)
        out = np.empty(int(np.prod(shape)))
        if self.L.orc_get_field(self.h, fid, _p(out)) != 0:
            raise ValueError(f"unknown field {fid}")
        return out.reshape(shape, order="F")

    def get_range(self, fid, lo, hi):
        lo_a = (ctypes.c_int * 3)(*lo)
        hi_a = (ctypes.c_int * 3)(*hi)
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        out = np.empty(int(np.prod(shape)))
        if self.L.orc_get_field_range(self.h, fid, lo_a, hi_a, _p(out)) != 0:
            raise ValueError(f"unknown field {fid}")
        return out.reshape(shape, order="F")

    def set_range(self, fid, lo, hi, values):
        lo_a = (ctypes.c_int * 3)(*lo)
        hi_a = (ctypes.c_int * 3)(*hi)
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        buf = _f64(np.broadcast_to(np.asarray(values, dtype=np.float64), shape))
        if self.L.orc_set_field_range(self.h, fid, lo_a, hi_a, _p(buf)) != 0:
            raise ValueError(f"unknown field {fid}")

    def fill_halos(self, fill_normal=True):
        self.L.orc_fill_halos(self.h, 1 if fill_normal else 0)

    def update_state(self):
        self.L.orc_update_state(self.h)

    def compute_closure(self):
        self.L.orc_compute_closure(self.h)

    def compute_hydrostatic(self):
        self.L.orc_compute_hydrostatic(self.h)

    def compute_tendencies(self):
        self.L.orc_compute_tendencies(self.h)

    def set_finalize(self, enforce_incompressibility=True):
        self.L.orc_set_finalize(self.h, 1 if enforce_incompressibility else 0)

    def time_step(self, dt, n=1):
        self.L.orc_time_step(self.h, float(dt), int(n))

    def rk3_substep(self, dt, gamma, zeta, first):
        self.L.orc_rk3_substep(self.h, float(dt), float(gamma), float(zeta), 1 if first else 0)

    def pressure_correction(self, dt):
        self.L.orc_pressure_correction(self.h, float(dt))

    def clock(self):
        t = ctypes.c_double()
        it = ctypes.c_int64()
        st = ctypes.c_int()
        self.L.orc_get_clock(self.h, ctypes.byref(t), ctypes.byref(it), ctypes.byref(st))
        return t.value, it.value, st.value

    def set_clock(self, time, iteration, stage=1, last_dt=float("inf"), last_stage_dt=float("inf")):
        self.L.orc_set_clock(self.h, float(time), int(iteration), int(stage), float(last_dt), float(last_stage_dt))

    def load_state(self, state, names):
        """restore_prognostic_state! from a checkpoint dictionary {group path: parent array} (the
        reference's JLD2 group names): prognostic fields, Gⁿ, G⁻, pNHS and the clock."""
        g = self.grid
        lo = tuple(1 - h for h in g.H)
        for q, n in enumerate(names):
            group = "velocities" if q < 3 else "tracers"
            for key, fid in ((f"{group}/{n}/data", q), (f"timestepper/Gⁿ/{n}/data", 100 + q),
                             (f"timestepper/G⁻/{n}/data", 200 + q)):
                a = np.asarray(state[key])
                self.set_range(fid, lo, tuple(l + s - 1 for l, s in zip(lo, a.shape)), a)
        p = np.asarray(state["pressures/pNHS/data"])
        self.set_range(300, lo, tuple(l + s - 1 for l, s in zip(lo, p.shape)), p)
        self.set_clock(state["clock/time"], state["clock/iteration"], state["clock/stage"],
                       state["clock/last_Δt"], state["clock/last_stage_Δt"])

    def divergence(self):
        g = self.grid
        out = np.empty(g.Nx * g.Ny * g.Nz)
        self.L.orc_divergence(self.h, _p(out))
        return out.reshape((g.Nx, g.Ny, g.Nz), order="F")

    def poisson_solve(self, rhs):
        g = self.grid
        r = _f64(rhs)
        out = np.empty_like(r)
        self.L.orc_poisson_solve(self.h, _p(r), _p(out))
        return out.reshape((g.Nx, g.Ny, g.Nz), order="F")

    def last_solve_info(self):
        """(iterations, recursive residual norm, recomputed FP64 residual norm) of the last CG solve."""
        it, r, t = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self.L.orc_last_solve_info(self.h, ctypes.byref(it), ctypes.byref(r), ctypes.byref(t))
        return it.value, r.value, t.value

    def cell_advection_timescale(self):
        return self.L.orc_cell_advection_timescale(self.h)

    def cg_precondition(self, kind, r):
        """z = M⁻¹ r for the CG preconditioner `kind` (OM_PRECOND_*), r and z interior (Nx, Ny, Nz)."""
        g = self.grid
        rr = _f64(r)
        out = np.empty_like(rr)
        self.L.orc_cg_precondition(self.h, int(kind), _p(rr), _p(out))
        return out.reshape((g.Nx, g.Ny, g.Nz), order="F")

    def grid_z(self):
        g = self.grid
        n = g.Nz + 2 * g.H[2]
        a = [np.empty(n) for _ in range(4)]
        self.L.orc_grid_z(self.h, *[_p(x) for x in a])
        return a   # dzC, dzF, zC, zF for k = 1-Hz .. Nz+Hz


def set_threads(n):
    """OpenMP thread count of the oracle's later parallel regions (called from this thread)."""
    lib().orc_set_threads(int(n))
    return lib().orc_max_threads()


def weno5(S, left=True, fp32=False):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_weno5(_p(S), 1 if left else 0, 1 if fp32 else 0)


def centered4(S, fp32=False):
    """Centered(4) interpolation at face i of S = (ψ[i-2], ψ[i-1], ψ[i], ψ[i+1])."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_centered4(_p(S), 1 if fp32 else 0)


def weno3(S, left=True, fp32=False):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().orc_weno3(_p(S), 1 if left else 0, 1 if fp32 else 0)


def weno5_smoothness(S, left=True, fp32=False):
    S = np.ascontiguousarray(S, dtype=np.float64)
    b, w = np.empty(3), np.empty(3)
    lib().orc_weno5_smoothness(_p(S), 1 if left else 0, 1 if fp32 else 0, _p(b), _p(w))
    return b, w


def coefficients(fp32=False):
    w5, w3, c4, c2 = np.empty(9), np.empty(4), np.empty(4), np.empty(2)
    lib().orc_coefficients(1 if fp32 else 0, _p(w5), _p(w3), _p(c4), _p(c2))
    return w5.reshape(3, 3), w3.reshape(2, 2), c4, c2


def thomas(a, b, c, f):
    """Batched Thomas solve; a, c: (nbatch, n-1), b, f: (nbatch, n)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    nb, n = b.shape
    a = np.ascontiguousarray(np.broadcast_to(a, (nb, n - 1)), dtype=np.float64)
    c = np.ascontiguousarray(np.broadcast_to(c, (nb, n - 1)), dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    x = np.empty_like(f)
    lib().orc_thomas(n, nb, _p(a), _p(b), _p(c), _p(f), _p(x))
    return x

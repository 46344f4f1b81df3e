/*
 * oceananigans_mi355x.h — C ABI of the MI355X-native NonhydrostaticModel time step.
 *
 * One opaque handle per model.  All device memory is owned by the library;
 * host buffers passed in are borrowed for the duration of the call only.
 * Every entry point returns 0 on success and a negative OM_E* code on error;
 * om_last_error(handle) gives the message.  No C++ exception crosses the ABI.
 * One host thread per handle; all work is ordered on one HIP stream owned by
 * the handle and every call is synchronous on return.
 *
 * Reference interfaces replaced (paths relative to the Oceananigans.jl tree):
 *   om_create            NonhydrostaticModel(grid; advection, closure, buoyancy, coriolis,
 *                        tracers, boundary_conditions)
 *                        src/Models/NonhydrostaticModels/nonhydrostatic_model.jl:124-313
 *   om_destroy           finalizer of the model (Julia GC); no reference counterpart
 *   om_field_info        Field data / OffsetArray parent layout
 *                        src/Grids/new_data.jl:11-17, src/Grids/grid_utils.jl:74-80
 *   om_set_field         set!(field, array) src/Fields/set!.jl:127-146
 *   om_get_field         interior(field) / parent(field) copy to host
 *   om_fill_halos        fill_halo_regions!(fields; fill_normal_flow_bcs)
 *                        src/BoundaryConditions/fill_halo_regions.jl:20-38
 *   om_update_state      update_state!(model)
 *                        src/Models/NonhydrostaticModels/update_nonhydrostatic_model_state.jl:22-62
 *   om_compute_tendencies compute_tendencies!(model)
 *                        src/Models/NonhydrostaticModels/compute_nonhydrostatic_tendencies.jl:12-40
 *   om_rk3_substep       rk3_substep!(model, Δt, γ, ζ) → pressure_correction_rk3_substep!
 *                        src/Models/NonhydrostaticModels/nonhydrostatic_rk3_substep.jl:31-63
 *   om_cache_previous_tendencies cache_previous_tendencies!(model)
 *                        src/Models/NonhydrostaticModels/cache_nonhydrostatic_tendencies.jl:21-31
 *   om_pressure_correction compute_pressure_correction!(model, Δt) + make_pressure_correction!(model, Δt)
 *                        src/Models/NonhydrostaticModels/pressure_correction.jl:6-20,82-106
 *   om_time_step         time_step!(model::AbstractModel{<:RungeKutta3TimeStepper}, Δt)
 *                        src/TimeSteppers/runge_kutta_3.jl:103-168
 *   om_get_clock/om_set_clock  model.clock (time, iteration, stage, last_Δt)
 *                        src/TimeSteppers/clock.jl:147-173
 *   om_poisson_solve     solve!(ϕ, ::FFTBasedPoissonSolver / ::FourierTridiagonalPoissonSolver, b)
 *                        src/Solvers/fft_based_poisson_solver.jl:94-124,
 *                        src/Solvers/fourier_tridiagonal_poisson_solver.jl:231-260
 *   om_divergence        divᶜᶜᶜ(i, j, k, grid, u, v, w) src/Operators/divergence_operators.jl:16-19
 */
#ifndef OCEANANIGANS_MI355X_H
#define OCEANANIGANS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OM_ABI_VERSION 2

/* status codes */
#define OM_OK            0
#define OM_EINVAL       -1   /* invalid argument or unsupported configuration */
#define OM_EHIP         -2   /* HIP runtime failure; the handle is unusable afterwards */
#define OM_ENOMEM       -3
#define OM_ESTATE       -4   /* call not valid in the current state */
#define OM_ENOTCONV     -5   /* a conjugate-gradient solve of the call stopped at maxiter above its tolerance;
                                 the call's work completed, as the reference's solver stops there and the
                                 step goes on (conjugate_gradient_solver.jl:279-284); the handle stays usable
                                 and om_last_solve_info describes the last solve */

/* topology per dimension (src/Grids/topology ... Periodic, Bounded, Flat) */
#define OM_PERIODIC 0
#define OM_BOUNDED  1
#define OM_FLAT     2

/* advection scheme (flux form) */
#define OM_ADV_NONE       0
#define OM_ADV_CENTERED2  1   /* Centered(order=2) */
#define OM_ADV_WENO5      2   /* WENO(order=5): buffer WENO(order=3) → Centered(order=2), advecting velocity Centered(order=4) */

/* turbulence closure */
#define OM_CLOSURE_NONE                 0
#define OM_CLOSURE_AMD                  1  /* AnisotropicMinimumDissipation(C=Cν=Cκ), Cb = nothing   */
#define OM_CLOSURE_SMAGORINSKY_LILLY    2  /* SmagorinskyLilly(C, Cb, Pr)                             */
#define OM_CLOSURE_SCALAR_DIFFUSIVITY   3  /* ScalarDiffusivity(ν, κ) (isotropic, constant)            */

/* buoyancy */
#define OM_BUOYANCY_NONE        0
#define OM_BUOYANCY_TRACER      1  /* BuoyancyTracer, tracer index buoyancy_tracer              */
#define OM_BUOYANCY_LINEAR_EOS  2  /* SeawaterBuoyancy(LinearEquationOfState(α, β)), tracers T,S */

/* boundary condition kinds on the bottom/top of a Bounded z */
#define OM_BC_DEFAULT   0  /* NoFlux for Center fields, Impenetrable for w */
#define OM_BC_FLUX      1
#define OM_BC_VALUE     2
#define OM_BC_GRADIENT  3

/* pressure solver */
#define OM_SOLVER_FFT                  0
#define OM_SOLVER_FOURIER_TRIDIAGONAL  1
#define OM_SOLVER_CONJUGATE_GRADIENT   2   /* ConjugateGradientPoissonSolver */

/* preconditioner of the conjugate-gradient Poisson solver */
#define OM_PRECOND_DEFAULT              0  /* the reference's default: DiagonallyDominantPreconditioner
                                              (FFT-based on an immersed grid whose underlying grid has one) */
#define OM_PRECOND_NONE                 1
#define OM_PRECOND_DIAGONALLY_DOMINANT  2
#define OM_PRECOND_FFT                  3  /* FFT-based solver of the underlying grid (Fourier–tridiagonal on stretched z) */
#define OM_PRECOND_COLUMNWISE_TRIDIAGONAL 4 /* ColumnwiseTridiagonalPreconditioner: per-column vertical part of V∇²,
                                               conjugate_gradient_poisson_solver.jl:298-397 */
#define OM_PRECOND_LEARNED              5  /* the FFT preconditioner plus a learned convolutional correction,
                                              in flexible CG with an FFT fallback (no reference counterpart;
                                              it replaces the preconditioner call of
                                              conjugate_gradient_poisson_solver.jl:169-211); Nx a multiple
                                              of 128, Ny and Nz multiples of 8, Bounded z */

#define OM_MAX_TRACERS 8
#define OM_MAX_FIELDS  (3 + OM_MAX_TRACERS)

/* field identifiers for om_field_info / om_set_field / om_get_field */
#define OM_FIELD_U        0
#define OM_FIELD_V        1
#define OM_FIELD_W        2
#define OM_FIELD_TRACER0  3      /* tracer n is OM_FIELD_TRACER0 + n */
#define OM_FIELD_GN       100    /* Gⁿ of prognostic field f is OM_FIELD_GN + f */
#define OM_FIELD_GM       200    /* G⁻ of prognostic field f is OM_FIELD_GM + f */
#define OM_FIELD_PNHS     300
#define OM_FIELD_PHY      301    /* hydrostatic pressure anomaly pHY′ */
#define OM_FIELD_NU       302    /* eddy viscosity νₑ */
#define OM_FIELD_KAPPA0   303    /* eddy diffusivity κₑ of tracer n is OM_FIELD_KAPPA0 + n (AMD) */
#define OM_FIELD_RHS      400    /* Poisson right-hand side / solution scratch (Center, interior only) */

typedef struct {
    int32_t precision;            /* 4 = FP32, 8 = FP64 (fixed at create) */
    int32_t N[3];                 /* Nx, Ny, Nz */
    int32_t H[3];                 /* halo widths (0 on a Flat dimension) */
    int32_t topology[3];          /* OM_PERIODIC / OM_BOUNDED / OM_FLAT */
    double  origin[3];            /* position of the first interior face in x, y, z */
    double  extent[3];            /* Lx, Ly, Lz */
    const double *z_faces;        /* NULL: uniform z; else Nz+1 increasing interior z faces */
    int32_t advection;            /* OM_ADV_* */
    int32_t closure;              /* OM_CLOSURE_* */
    double  closure_params[4];    /* AMD: {Cν, Cκ}; SmagorinskyLilly: {C, Cb, Pr}; ScalarDiffusivity: {ν, κ} */
    int32_t buoyancy;             /* OM_BUOYANCY_* */
    int32_t buoyancy_tracer;      /* tracer index of b (BuoyancyTracer) */
    int32_t T_tracer, S_tracer;   /* tracer indices of T and S (linear EOS) */
    double  gravitational_acceleration;
    double  thermal_expansion, haline_contraction;
    int32_t coriolis;             /* 0 none, 1 FPlane */
    double  f;                    /* Coriolis parameter of the FPlane */
    int32_t n_tracers;
    int32_t bc_bottom[OM_MAX_FIELDS];      /* OM_BC_* per prognostic field (u, v, w, tracers...) */
    int32_t bc_top[OM_MAX_FIELDS];
    double  bc_bottom_value[OM_MAX_FIELDS];
    double  bc_top_value[OM_MAX_FIELDS];
    int32_t pressure_solver;      /* OM_SOLVER_*; stretched z requires FOURIER_TRIDIAGONAL or CONJUGATE_GRADIENT */
    int32_t hydrostatic_pressure; /* -1: reference default rule; 0 off; 1 on */
    /* ConjugateGradientPoissonSolver (conjugate_gradient_poisson_solver.jl:134-163) */
    double  cg_reltol;            /* < 0: sqrt(eps(FT)) */
    double  cg_abstol;            /* < 0: sqrt(eps(FT)) */
    int32_t cg_maxiter;           /* ≤ 0: Nx·Ny·Nz */
    int32_t cg_preconditioner;    /* OM_PRECOND_* */
    /* GridFittedBottom immersed boundary (grid_fitted_bottom.jl:17-20,118-147): bottom height z_b(x, y)
     * at the Nx·Ny cell centres (x fastest), or NULL for none.  Cells whose centre lies at or below
     * z_b are immersed (CenterImmersedCondition).  Requires WENO(order=5) or no advection. */
    const double *bottom_height;
    int32_t reserved[8];
} om_config_t;

typedef struct {
    void    *data;          /* device pointer to parent(field)[1,1,1] (first halo point) */
    int32_t  element_size;  /* 4 or 8 */
    int32_t  parent_size[3];/* Nx+2Hx(+1 on a Bounded Face axis), ... */
    int32_t  offset[3];     /* index of the first parent element in reference indices (1 - H) */
    int64_t  stride[3];     /* element strides (x fastest: 1, nx, nx*ny) */
} om_field_info_t;

typedef struct om_model om_model_t;

/* lifecycle */
int         om_create(const om_config_t *cfg, om_model_t **out);
int         om_destroy(om_model_t *m);
const char *om_last_error(const om_model_t *m);   /* per handle; om_last_error(NULL) gives the create error */
int         om_abi_version(void);

/* data movement (host buffers are FP64, x fastest) */
int om_field_info(om_model_t *m, int32_t field, om_field_info_t *info);
int om_set_field(om_model_t *m, int32_t field, const double *host, int32_t with_halos);
int om_get_field(om_model_t *m, int32_t field, double *host, int32_t with_halos);

/* time stepping (reference semantics) */
int om_fill_halos(om_model_t *m, uint32_t field_mask, int32_t fill_normal_flow);
int om_update_state(om_model_t *m);
int om_compute_tendencies(om_model_t *m);
int om_rk3_substep(om_model_t *m, double dt, double gamma, double zeta, int32_t first_stage);
int om_cache_previous_tendencies(om_model_t *m);
int om_pressure_correction(om_model_t *m, double dt);
int om_time_step(om_model_t *m, double dt, int32_t nsteps);

/* clock */
int om_get_clock(const om_model_t *m, double *time, int64_t *iteration, int32_t *stage,
                 double *last_dt, double *last_stage_dt);
int om_set_clock(om_model_t *m, double time, int64_t iteration, int32_t stage,
                 double last_dt, double last_stage_dt);

/* solver building blocks */
int om_poisson_solve(om_model_t *m, const double *rhs_interior, double *phi_interior);
/* the last conjugate-gradient solve: iterations, the recursively updated residual norm ‖V⁻¹r‖ the
 * iteration stopped on, and ‖V⁻¹(e − ē)‖ with e = b − A x recomputed in FP64 from the final x, in the
 * iteration's zero-mean gauge (conjugate_gradient_solver.jl:177-284; no reference counterpart for the
 * recomputed residual) */
int om_last_solve_info(const om_model_t *m, int32_t *iterations, double *residual_norm, double *true_residual_norm);
int om_divergence(om_model_t *m, double *div_interior);

/* diagnostics reductions on the device
 * om_cell_advection_timescale: cell_advection_timescale(model)
 *   (src/Models/NonhydrostaticModels/NonhydrostaticModels.jl:93-97, src/Advection/cell_advection_timescale.jl:14-35):
 *   the minimum over interior cells of 1 / (|u|/Δxᶠᶜᶜ + |v|/Δyᶜᶠᶜ + |w|/Δzᶜᶜᶠ) in the model's precision;
 *   +Inf for a fluid at rest, NaN when a velocity is NaN.
 * om_field_statistics: maximum, minimum and the number of NaNs of a field's data, over its parent
 *   array (with_halos = 1: maximum(field.data.parent) in
 *   src/TurbulenceClosures/turbulence_closure_diagnostics.jl:50-70; any(isnan, parent(field)) in
 *   src/Diagnostics/nan_checker.jl:37) or its interior (with_halos = 0); NaNs are excluded from the
 *   maximum and minimum. */
int om_cell_advection_timescale(om_model_t *m, double *tau);

/* learned preconditioner (no reference counterpart; see OM_PRECOND_LEARNED).  The network: a
 * four-level hierarchy of seven 3×3×3 convolutions of 16 channels on the grid and on the grid
 * coarsened 2×, 4× and 8× (2×2×2 means down, copies up, skip sums; ReLU after all but the last),
 * BF16 weights and activations, FP32 accumulation on the matrix cores.  Needs Nx a multiple of 128,
 * Ny and Nz multiples of 8.  Input channels per cell: r V⁻¹ / rms, z_FFT / rms, active-cell mask;
 * output: the correction to z_FFT in units of rms(z_FFT).
 * Weights: OM_LEARNED_NWEIGHTS FP32 values, layout [layer][tap][cin][cout], tap = (dz+1)·9 +
 * (dy+1)·3 + (dx+1).  Host arrays of the test hooks are dense, x fastest, channels innermost.
 *   om_learned_set_weights / get_weights   the FP32 master weights (setting them resets Adam)
 *   om_learned_log       log (input, target) pairs from the first `iterations_per_solve`
 *                        iterations of every later FFT- or learned-preconditioned CG solve, into a
 *                        ring of `capacity` pairs on the device; target = (x_final − x_k)/rms − ẑ_k;
 *                        each call starts a new set
 *   om_learned_pairs     number of pairs held
 *   om_learned_train     `steps` Adam steps (lr) over the held pairs, cycling; loss of the first
 *                        and of the last step: the A-energy of the remaining error relative to
 *                        the target's, (δ − t)ᵀ(−V∇²)(δ − t) / tᵀ(−V∇²)t, with the solver's operator
 *   om_learned_evaluate  that loss averaged over the held pairs, without training (1 = the FFT
 *                        preconditioner alone, 0 = the exact inverse)
 *   om_learned_forward   network output for a given 3-channel input (test hook)
 *   om_learned_backward  mean-square loss over the mask, weight gradients and input gradient
 *                        (3 channels) for a given input, target and mask (test hook)
 *   om_learned_last_solve  iterations that used the network in the last solve, and whether the
 *                        solve handed over to the FFT preconditioner
 *   om_learned_energy_loss  the training loss's energy and its gradient for a given network output
 *                        y and target t (test hook): d = active·(y − t), E = dᵀ(−V∇²)d with the
 *                        CG solver's conditional operator, dE/dy = −2·active·V∇²d (dense, x fastest) */
#define OM_LEARNED_NWEIGHTS (7 * 27 * 16 * 16)
int om_learned_set_weights(om_model_t *m, const float *weights);
int om_learned_get_weights(om_model_t *m, float *weights);
int om_learned_log(om_model_t *m, int32_t iterations_per_solve, int32_t capacity);
int om_learned_pairs(om_model_t *m, int32_t *count);
int om_learned_pairs_logged(om_model_t *m, int64_t *total);   /* pairs logged since om_learned_log (the ring keeps the last `capacity`) */
int om_learned_train(om_model_t *m, int32_t steps, double lr, double *loss_first, double *loss_last);
int om_learned_evaluate(om_model_t *m, double *mean_loss);
int om_learned_forward(om_model_t *m, const float *input3, float *output);
int om_learned_backward(om_model_t *m, const float *input3, const float *target, const float *mask, double *loss,
                        float *dweights, float *dinput3);
int om_learned_last_solve(om_model_t *m, int32_t *learned_iterations, int32_t *fell_back);
int om_learned_energy_loss(om_model_t *m, const float *output, const float *target, double *energy, float *doutput);
int om_field_statistics(om_model_t *m, int32_t field, int32_t with_halos, double *maximum, double *minimum,
                        int64_t *nan_count);
/* om_field_statistics_noalloc: the same statistics as om_field_statistics, computed without any
 * allocation.  No counterpart in the reference: it is for callers that cannot allocate device or
 * pinned memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its reduction works in a file-scope device array of the library, 3·1024 + 3 doubles
 * (24 KiB) held by each GPU the library runs on until the process ends, instead of the handle's
 * buffers; the result is copied to the caller's host variables before the call returns.  Fields of
 * any model size are accepted up to 2^40 elements (OM_EINVAL beyond).  Every call on one GPU uses
 * that GPU's one array: calls that overlap on one GPU (from several host threads) race. */
int om_field_statistics_noalloc(om_model_t *m, int32_t field, int32_t with_halos, double *maximum, double *minimum,
                                int64_t *nan_count);

/* timing of the last om_time_step call (HIP events on the handle's stream).
 * ms_total: the whole call.  With phase timing enabled, ms_per_phase[p] and launches_per_phase[p]
 * (p < nphases ≤ OM_NPHASES) give the summed device time and the number of timed launches of
 * each phase below.  No reference counterpart (the reference times whole steps only,
 * benchmarking/src/utils.jl:84-101). */
#define OM_PHASE_RK_UPDATE    0   /* RK3 update + flux-BC tendencies                        */
#define OM_PHASE_POISSON      1   /* source term + transforms + eigenvalue / tridiagonal solve */
#define OM_PHASE_CORRECTION   2   /* U -= ∇p, p /= Δt                                        */
#define OM_PHASE_HALOS        3   /* halo fills                                              */
#define OM_PHASE_CLOSURE      4   /* eddy viscosity / diffusivity                            */
#define OM_PHASE_HYDROSTATIC  5   /* pHY′ column integral                                    */
#define OM_PHASE_TENDENCIES   6   /* fused tendency kernel                                   */
#define OM_NPHASES            7
int om_set_phase_timing(om_model_t *m, int32_t enable);
int om_last_step_timing(const om_model_t *m, double *ms_total, double *ms_per_phase, int32_t *launches_per_phase,
                        int32_t nphases);

/* launch-shape choices of this library (no reference counterpart).  Results do not depend on
 * them: every tendency kernel gives identically rounded tendencies. */
#define OM_TUNE_TENDENCY_KERNEL 0   /* 0: z-marching LDS-tiled tendency and closure kernels where supported
                                       (default); 1: thread-per-cell kernels */
#define OM_TUNE_KCHUNK          1   /* z levels marched by one block of the tiled kernels (default 16;
                                       om_autotune may choose different ones for the two kernels) */
#define OM_TUNE_LEARNED_ITERATIONS 2 /* leading iterations of each CG solve that apply the learned
                                        correction (0: all); later iterations use the FFT
                                        preconditioner.  Solutions agree to the solver tolerance. */
#define OM_TUNE_POISSON_REFINEMENT 3 /* 1: each FFT / Fourier-tridiagonal Poisson solve is followed by
                                        one step of iterative refinement: the residual b - Lp of the
                                        computed solution in FP64 arithmetic (for om_poisson_solve
                                        against the FP64 right-hand side), one more solve for the
                                        correction, p += correction.  0 (default): off.  Not with
                                        the conjugate-gradient solver (its residuals are FP64
                                        already) or an immersed bottom. */
#define OM_TUNE_STEP_GRAPHS 4      /* 1 (default): om_time_step advances pairs of steps by replaying the
                                       model's two-step launch sequence captured once into a hipGraph
                                       (the same launches; recaptured after a change of Δt or of a
                                       launch choice); 0: every launch issued directly.  Phase timing
                                       and the conjugate-gradient solver always launch directly. */
#define OM_TUNE_CONV_TILED 5       /* convolutions of the learned preconditioner: 1 (default) input tiles
                                       staged in LDS; 0 the input read per tap.  Identical results. */
#define OM_TUNE_WGRAD_MFMA 6       /* weight gradients of the learned preconditioner: 1 (default) on the
                                       matrix cores; 0 FP32 sums.  Equal to FP32 rounding. */
int om_set_tuning(om_model_t *m, int32_t key, int32_t value);
int om_get_tuning(const om_model_t *m, int32_t key, int32_t *value);

/* Autotuner (no reference counterpart): times the tiled tendency and closure kernels, each on
 * its own, on the model's current state for z-chunk lengths 4 .. 64 and keeps the fastest for
 * each (OM_TUNE_KCHUNK sets both).  Results do not depend on the choice.  *kchunk (may be NULL)
 * receives the tendency kernel's. */
int om_autotune(om_model_t *m, int32_t *kchunk);

/* Ensembles of independent models (no reference counterpart for NonhydrostaticModel: the
 * reference advances such members one after another, `time_step!` per model).  Members are
 * created with om_create, may differ in size and configuration, and stay owned by the caller;
 * the conjugate-gradient solver is not allowed (its steps read scalars on the host).
 * om_ensemble_time_step advances every member nsteps steps of Δt; `use_graphs` picks the form
 * (OM_ENSEMBLE_*): member graphs (1) record each member's two-step launch sequence once into a
 * hipGraph on its own stream and replay it, members issued in rounds so that their streams run
 * concurrently; direct (0) issues the same launches member by member; batched (2) records every
 * member's sequence, merges position p of all sequences into one launch over a device table of the
 * members' arguments (each member's blocks a contiguous range of one grid), captures the merged
 * sequence into one hipGraph on the ensemble's stream and replays it per pair of steps.  Members at
 * iteration 0 take their first step on their own first.  It returns when every member is done. */
#define OM_ENSEMBLE_DIRECT        0   /* every member's launches issued directly, member by member      */
#define OM_ENSEMBLE_MEMBER_GRAPHS 1   /* each member's two-step sequence replayed as its own hipGraph    */
#define OM_ENSEMBLE_BATCHED       2   /* one launch per kernel of the step covering every member (member
                                         tables), the whole two-step sequence one hipGraph; members must
                                         share one launch sequence (same precision, closure, buoyancy,
                                         tracer count and an FFT solver on a uniform grid; sizes may differ) */
typedef struct om_ensemble om_ensemble_t;
int         om_ensemble_create(om_model_t *const *members, int32_t n, om_ensemble_t **out);
int         om_ensemble_time_step(om_ensemble_t *e, double dt, int32_t nsteps, int32_t use_graphs);
int         om_ensemble_destroy(om_ensemble_t *e);
const char *om_ensemble_last_error(const om_ensemble_t *e);   /* NULL: the last create error */
/* the batched form's member groups (1 .. 64, default 1): members are split in order into groups of
 * about equal work, each group's member-table graph on its own stream, the groups' graphs running
 * concurrently (no reference counterpart) */
int         om_ensemble_set_batch_groups(om_ensemble_t *e, int32_t groups);

#ifdef __cplusplus
}
#endif

#endif /* OCEANANIGANS_MI355X_H */

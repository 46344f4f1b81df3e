# OceananigansMI355XExt.jl — Julia package extension that routes the NonhydrostaticModel RK3 time
# step on a RectilinearGrid (optionally with a GridFittedBottom) to liboceananigans_mi355x.so
# through its C ABI (include/oceananigans_mi355x.h).
#
# Seams used (paths relative to the Oceananigans.jl tree):
#   architecture      src/Architectures.jl:21-113 (AbstractSerialArchitecture; the method list of
#                     ext/OceananigansAMDGPUExt.jl:36-111)
#   pressure solver   src/Models/NonhydrostaticModels/NonhydrostaticModels.jl:40-76
#   time stepping     src/TimeSteppers/runge_kutta_3.jl:103-168 (time_step!)
#   state update      src/Models/NonhydrostaticModels/update_nonhydrostatic_model_state.jl:22-62
#   set!              src/Models/NonhydrostaticModels/set_nonhydrostatic_model.jl:39-74
# The Reactant extension overrides the same seams (ext/OceananigansReactantExt/TimeSteppers.jl:22-30).
#
# Storage.  The library owns the model's device state (fields, tendencies, pressures, the solver).
# The Julia model's own Fields live on the host (array_type(::MI355X) = Array) and mirror it:
# set!(model; ...) writes the host fields and pushes their interiors to the device; time_step!
# advances the device state and, when the architecture was built with sync = true (the default),
# copies the prognostic fields and the pressures back into the host mirrors, so that output
# writers, diagnostics and user code read current values.  With sync = false the mirrors are
# refreshed by pull!(model).
module OceananigansMI355XExt

using Oceananigans
using Oceananigans.Architectures: AbstractSerialArchitecture, CPU
using Oceananigans.Grids: RectilinearGrid, Periodic, Bounded, Flat, NegativeZDirection, topology
using Oceananigans.ImmersedBoundaries: ImmersedBoundaryGrid, GridFittedBottom
using Oceananigans.Advection: WENO, Centered, FluxFormAdvection
using Oceananigans.TurbulenceClosures: AnisotropicMinimumDissipation, ScalarDiffusivity, Smagorinsky,
                                       LillyCoefficient, ThreeDimensionalFormulation
using Oceananigans.BuoyancyFormulations: BuoyancyForce, BuoyancyTracer, SeawaterBuoyancy, LinearEquationOfState
using Oceananigans.Coriolis: FPlane
using Oceananigans.BoundaryConditions: BoundaryCondition, Flux, Value, Gradient, NormalFlow
using Oceananigans.Solvers: FFTBasedPoissonSolver, FourierTridiagonalPoissonSolver, ConjugateGradientPoissonSolver,
                            DiagonallyDominantPreconditioner, ColumnwiseTridiagonalPreconditioner
using Oceananigans.TimeSteppers: RungeKutta3TimeStepper
using Oceananigans.Fields: interior
using KernelAbstractions: KernelAbstractions as KA

import Oceananigans.Architectures as AC
import Oceananigans.BoundaryConditions as BC
import Oceananigans.Fields as FD
import Oceananigans.Utils as UT
import Oceananigans.Models.NonhydrostaticModels as NM
import Oceananigans.TimeSteppers: time_step!, update_state!

const libom = get(ENV, "OCEANANIGANS_MI355X_LIB", "liboceananigans_mi355x")

# --- constants of include/oceananigans_mi355x.h ---------------------------------------------------
const OM_MAX_TRACERS = 8
const OM_MAX_FIELDS = 3 + OM_MAX_TRACERS
const OM_PERIODIC, OM_BOUNDED, OM_FLAT = Int32(0), Int32(1), Int32(2)
const OM_ADV_NONE, OM_ADV_CENTERED2, OM_ADV_WENO5 = Int32(0), Int32(1), Int32(2)
const OM_CLOSURE_NONE, OM_CLOSURE_AMD, OM_CLOSURE_SMAGORINSKY_LILLY, OM_CLOSURE_SCALAR_DIFFUSIVITY = Int32.((0, 1, 2, 3))
const OM_BUOYANCY_NONE, OM_BUOYANCY_TRACER, OM_BUOYANCY_LINEAR_EOS = Int32.((0, 1, 2))
const OM_BC_DEFAULT, OM_BC_FLUX, OM_BC_VALUE, OM_BC_GRADIENT = Int32.((0, 1, 2, 3))
const OM_SOLVER_FFT, OM_SOLVER_FOURIER_TRIDIAGONAL, OM_SOLVER_CONJUGATE_GRADIENT = Int32.((0, 1, 2))
const OM_PRECOND_DEFAULT, OM_PRECOND_NONE, OM_PRECOND_DIAGONALLY_DOMINANT, OM_PRECOND_FFT,
      OM_PRECOND_COLUMNWISE_TRIDIAGONAL = Int32.((0, 1, 2, 3, 4))
const OM_FIELD_TRACER0, OM_FIELD_PNHS, OM_FIELD_PHY = Int32(3), Int32(300), Int32(301)

# --- architecture ------------------------------------------------------------------------------
"""
    MI355X(; device = 0, sync = true)

One AMD Instinct MI355X, driven through liboceananigans_mi355x.  `device` is the HIP device the
library runs on (the one current in the calling thread; select it with ROCR_VISIBLE_DEVICES).
`sync = true` copies the prognostic fields and the pressures back into the model's host Fields
after every `time_step!`.
"""
struct MI355X <: AbstractSerialArchitecture
    device :: Int
    sync :: Bool
end
MI355X(; device = 0, sync = true) = MI355X(device, sync)

Base.summary(::MI355X) = "MI355X"

# the method list of ext/OceananigansAMDGPUExt.jl:36-111, for host-mirrored storage
AC.device(::MI355X) = KA.CPU()                       # host-side kernels (set!, diagnostics) run on the CPU
AC.device!(arch::MI355X, i) = i == arch.device ? nothing :
    throw(ArgumentError("MI355X: the library runs on the device current in the calling thread; select it with ROCR_VISIBLE_DEVICES"))
AC.ndevices(::MI355X) = 1
AC.synchronize(::MI355X) = nothing                   # every library call is synchronous on return
UT.sync_device!(::MI355X) = nothing
AC.array_type(::MI355X) = Array
AC.cpu_architecture(::MI355X) = CPU()
AC.child_architecture(arch::MI355X) = arch
AC.on_architecture(::MI355X, a::Number) = a
AC.on_architecture(::MI355X, a::Array) = a
AC.on_architecture(::MI355X, a::BitArray) = a
AC.on_architecture(::MI355X, a::StepRangeLen) = a
AC.unified_array(::MI355X, a::AbstractArray) = a
BC.validate_boundary_condition_architecture(::Array, ::MI355X, bc, side) = nothing
# set!(::Field, ::Array) is the reference's own method for host arrays (src/Fields/set!.jl:127-146);
# the host Fields of an MI355X model are Arrays, so it applies unchanged.

# The pressure solve runs in the library: the model holds a descriptor of the solver kind instead
# of a host solver (NonhydrostaticModels.jl:40-76 picks FFT for a regular grid, Fourier–tridiagonal
# for a stretched z, and the FFT solver of the underlying grid for an immersed one).
struct MI355XPoissonSolver
    kind :: Int32
end
regular_z(grid::RectilinearGrid) = grid.z.Δᵃᵃᶠ isa Number
NM.nonhydrostatic_pressure_solver(::MI355X, grid::RectilinearGrid, ::Nothing) =
    MI355XPoissonSolver(regular_z(grid) ? OM_SOLVER_FFT : OM_SOLVER_FOURIER_TRIDIAGONAL)
NM.nonhydrostatic_pressure_solver(arch::MI355X, ibg::ImmersedBoundaryGrid, ::Nothing) =
    NM.nonhydrostatic_pressure_solver(arch, ibg.underlying_grid, nothing)

# --- C structs (field order and types follow include/oceananigans_mi355x.h; tests/test_abi.py parses
# this declaration and checks it against the header) ----------------------------------------------
struct OmConfig
    precision::Int32
    N::NTuple{3,Int32}
    H::NTuple{3,Int32}
    topology::NTuple{3,Int32}
    origin::NTuple{3,Float64}
    extent::NTuple{3,Float64}
    z_faces::Ptr{Float64}
    advection::Int32
    closure::Int32
    closure_params::NTuple{4,Float64}
    buoyancy::Int32
    buoyancy_tracer::Int32
    T_tracer::Int32
    S_tracer::Int32
    gravitational_acceleration::Float64
    thermal_expansion::Float64
    haline_contraction::Float64
    coriolis::Int32
    f::Float64
    n_tracers::Int32
    bc_bottom::NTuple{OM_MAX_FIELDS,Int32}
    bc_top::NTuple{OM_MAX_FIELDS,Int32}
    bc_bottom_value::NTuple{OM_MAX_FIELDS,Float64}
    bc_top_value::NTuple{OM_MAX_FIELDS,Float64}
    pressure_solver::Int32
    hydrostatic_pressure::Int32
    cg_reltol::Float64
    cg_abstol::Float64
    cg_maxiter::Int32
    cg_preconditioner::Int32
    bottom_height::Ptr{Float64}
    reserved::NTuple{8,Int32}
end

struct OmFieldInfo
    data::Ptr{Cvoid}
    element_size::Int32
    parent_size::NTuple{3,Int32}
    offset::NTuple{3,Int32}
    stride::NTuple{3,Int64}
end

# --- thin ccall layer: one Julia function per C entry point ------------------------------------
last_error(h) = unsafe_string(ccall((:om_last_error, libom), Cstring, (Ptr{Cvoid},), h))
om_check(rc, h) = rc == 0 || throw(ErrorException("liboceananigans_mi355x: " * last_error(h)))

function om_create(cfg::OmConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:om_create, libom), Cint, (Ref{OmConfig}, Ref{Ptr{Cvoid}}), cfg, h)
    rc == 0 || throw(ArgumentError("liboceananigans_mi355x: " * last_error(C_NULL)))
    return h[]
end
om_destroy(h) = ccall((:om_destroy, libom), Cint, (Ptr{Cvoid},), h)
om_set_field!(h, id, a::Array{Float64,3}, with_halos) =
    om_check(ccall((:om_set_field, libom), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32), h, id, a, with_halos), h)
om_get_field!(a::Array{Float64,3}, h, id, with_halos) =
    om_check(ccall((:om_get_field, libom), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32), h, id, a, with_halos), h)
function om_field_info(h, id)
    info = Ref{OmFieldInfo}()
    om_check(ccall((:om_field_info, libom), Cint, (Ptr{Cvoid}, Int32, Ref{OmFieldInfo}), h, id, info), h)
    return info[]
end
om_fill_halos!(h, mask, fill_normal) =
    om_check(ccall((:om_fill_halos, libom), Cint, (Ptr{Cvoid}, UInt32, Int32), h, mask, fill_normal), h)
om_update_state!(h) = om_check(ccall((:om_update_state, libom), Cint, (Ptr{Cvoid},), h), h)
om_compute_tendencies!(h) = om_check(ccall((:om_compute_tendencies, libom), Cint, (Ptr{Cvoid},), h), h)
om_rk3_substep!(h, Δt, γ, ζ, first) =
    om_check(ccall((:om_rk3_substep, libom), Cint, (Ptr{Cvoid}, Float64, Float64, Float64, Int32), h, Δt, γ, ζ, first), h)
om_cache_previous_tendencies!(h) = om_check(ccall((:om_cache_previous_tendencies, libom), Cint, (Ptr{Cvoid},), h), h)
om_pressure_correction!(h, Δt) = om_check(ccall((:om_pressure_correction, libom), Cint, (Ptr{Cvoid}, Float64), h, Δt), h)
om_time_step!(h, Δt, n) = om_check(ccall((:om_time_step, libom), Cint, (Ptr{Cvoid}, Float64, Int32), h, Δt, n), h)
function om_get_clock(h)
    t = Ref{Float64}(); it = Ref{Int64}(); st = Ref{Int32}(); ldt = Ref{Float64}(); lsdt = Ref{Float64}()
    om_check(ccall((:om_get_clock, libom), Cint,
                   (Ptr{Cvoid}, Ref{Float64}, Ref{Int64}, Ref{Int32}, Ref{Float64}, Ref{Float64}), h, t, it, st, ldt, lsdt), h)
    return (time = t[], iteration = it[], stage = st[], last_Δt = ldt[], last_stage_Δt = lsdt[])
end
om_set_clock!(h, t, it, st, ldt, lsdt) =
    om_check(ccall((:om_set_clock, libom), Cint, (Ptr{Cvoid}, Float64, Int64, Int32, Float64, Float64), h, t, it, st, ldt, lsdt), h)

# --- om_config(model): the Julia twin of build_config in oceananigans_mi355x/models.py ----------
unsupported(what) = throw(ArgumentError("MI355X: $what is not supported on this architecture"))

topology_code(::Type{Periodic}) = OM_PERIODIC
topology_code(::Type{Bounded}) = OM_BOUNDED
topology_code(::Type{Flat}) = OM_FLAT
topology_code(T) = unsupported("topology $T")

ntuple_fill(v, n) = ntuple(_ -> v, n)
ntuple_set(t::NTuple{N,T}, i, v) where {N,T} = ntuple(q -> q == i ? T(v) : t[q], N)

# one scalar from a closure coefficient that may be given per tracer
function single_coefficient(c, what)
    c isa Number && return Float64(c)
    vals = unique(Float64.(values(c)))
    length(vals) == 1 || unsupported("per-tracer $what")
    return vals[1]
end

function advection_code(scheme)
    scheme === nothing && return OM_ADV_NONE
    if scheme isa FluxFormAdvection
        scheme.x == scheme.y == scheme.z || unsupported("direction-dependent advection")
        scheme = scheme.x
    end
    scheme isa WENO && Oceananigans.Advection.weno_order(scheme) == 5 && scheme.bounds === nothing && return OM_ADV_WENO5
    scheme isa Centered{1} && return OM_ADV_CENTERED2
    unsupported("advection scheme $(summary(scheme))")
end

# (kind, value) of a top or bottom condition; conditions must be constants
function bc_code(bc)
    bc === nothing && return OM_BC_DEFAULT, 0.0
    cond = bc.condition
    cond === nothing && (cond = 0)
    bc.classification isa NormalFlow && return OM_BC_DEFAULT, 0.0
    cond isa Number || unsupported("a non-constant boundary condition")
    bc.classification isa Flux && return OM_BC_FLUX, Float64(cond)
    bc.classification isa Value && return OM_BC_VALUE, Float64(cond)
    bc.classification isa Gradient && return OM_BC_GRADIENT, Float64(cond)
    unsupported("boundary condition $(summary(bc))")
end

function preconditioner_code(P)
    P === nothing && return OM_PRECOND_NONE
    P isa DiagonallyDominantPreconditioner && return OM_PRECOND_DIAGONALLY_DOMINANT
    P isa Union{FFTBasedPoissonSolver, FourierTridiagonalPoissonSolver} && return OM_PRECOND_FFT
    P isa ColumnwiseTridiagonalPreconditioner && return OM_PRECOND_COLUMNWISE_TRIDIAGONAL
    unsupported("preconditioner $(typeof(P))")
end

"""
    om_config(model) -> (OmConfig, keepalive)

Translate a NonhydrostaticModel on MI355X into the library's configuration.  `keepalive` holds the
arrays the struct points to (z faces, bottom height); they must outlive `om_create`.  Throws
ArgumentError for anything outside the supported set.
"""
function om_config(model)
    grid = model.grid
    keep = Any[]
    bottom = Ptr{Float64}(C_NULL)
    if grid isa ImmersedBoundaryGrid
        ib = grid.immersed_boundary
        ib isa GridFittedBottom || unsupported("immersed boundary $(typeof(ib))")
        grid = grid.underlying_grid
        hb = Array{Float64}(undef, grid.Nx, grid.Ny)
        zb = ib.bottom_height
        for j in 1:grid.Ny, i in 1:grid.Nx
            hb[i, j] = Float64(zb[i, j, 1])
        end
        push!(keep, hb)
        bottom = pointer(hb)
    end
    grid isa RectilinearGrid || unsupported("grid $(typeof(grid))")
    bg = model.background_fields
    all(x -> x isa Oceananigans.Fields.ZeroField, (values(bg.velocities)..., values(bg.tracers)...)) ||
        unsupported("background fields")
    model.particles === nothing || unsupported("Lagrangian particles")
    model.biogeochemistry === nothing || unsupported("biogeochemistry")
    model.stokes_drift === nothing || unsupported("Stokes drift")
    model.free_surface === nothing || unsupported("a free surface")
    all(f -> f === Oceananigans.Forcings.zeroforcing, values(model.forcing)) || unsupported("forcing")

    FT = eltype(grid)
    FT in (Float32, Float64) || unsupported("float type $FT")
    TX, TY, TZ = topology(grid)
    topo = (topology_code(TX), topology_code(TY), topology_code(TZ))
    N = Int32.((grid.Nx, grid.Ny, grid.Nz))
    H = Int32.((grid.Hx, grid.Hy, grid.Hz))
    origin = (TX === Flat ? 0.0 : Float64(grid.xᶠᵃᵃ[1]), TY === Flat ? 0.0 : Float64(grid.yᵃᶠᵃ[1]),
              TZ === Flat ? 0.0 : Float64(grid.z.cᵃᵃᶠ[1]))
    extent = (TX === Flat ? 1.0 : Float64(grid.Lx), TY === Flat ? 1.0 : Float64(grid.Ly), TZ === Flat ? 1.0 : Float64(grid.Lz))
    zfaces = Ptr{Float64}(C_NULL)
    if TZ !== Flat && !regular_z(grid)
        zf = Float64[grid.z.cᵃᵃᶠ[k] for k in 1:grid.Nz+1]
        push!(keep, zf)
        zfaces = pointer(zf)
    end

    names = keys(model.tracers)
    nt = length(names)
    nt <= OM_MAX_TRACERS || unsupported("more than $OM_MAX_TRACERS tracers")
    index(name) = (q = findfirst(==(name), names); q === nothing ? Int32(-1) : Int32(q - 1))

    closure = model.closure
    cl, cp = OM_CLOSURE_NONE, (0.0, 0.0, 0.0, 0.0)
    if closure isa AnisotropicMinimumDissipation
        closure.Cb === nothing || unsupported("AMD with a buoyancy modification Cb")
        cl, cp = OM_CLOSURE_AMD, (single_coefficient(closure.Cν, "Cν"), single_coefficient(closure.Cκ, "Cκ"), 0.0, 0.0)
    elseif closure isa Smagorinsky && closure.coefficient isa LillyCoefficient
        c = closure.coefficient
        cl, cp = OM_CLOSURE_SMAGORINSKY_LILLY, (Float64(c.smagorinsky), Float64(c.reduction_factor), single_coefficient(closure.Pr, "Pr"), 0.0)
    elseif closure isa ScalarDiffusivity{<:Any, ThreeDimensionalFormulation}
        cl, cp = OM_CLOSURE_SCALAR_DIFFUSIVITY, (single_coefficient(closure.ν, "ν"), single_coefficient(closure.κ, "κ"), 0.0, 0.0)
    elseif closure !== nothing
        unsupported("closure $(summary(closure))")
    end

    b, bt, Tt, St, g, α, β = OM_BUOYANCY_NONE, Int32(-1), Int32(-1), Int32(-1), 9.80665, 0.0, 0.0
    if model.buoyancy !== nothing
        bf = model.buoyancy
        bf isa BuoyancyForce && bf.gravity_unit_vector isa NegativeZDirection || unsupported("a tilted gravity vector")
        form = bf.formulation
        if form isa BuoyancyTracer
            b, bt = OM_BUOYANCY_TRACER, index(:b)
        elseif form isa SeawaterBuoyancy && form.equation_of_state isa LinearEquationOfState
            b, Tt, St = OM_BUOYANCY_LINEAR_EOS, index(:T), index(:S)
            g = Float64(form.gravitational_acceleration)
            α = Float64(form.equation_of_state.thermal_expansion)
            β = Float64(form.equation_of_state.haline_contraction)
        else
            unsupported("buoyancy formulation $(typeof(form))")
        end
    end

    co, f = Int32(0), 0.0
    if model.coriolis isa FPlane
        co, f = Int32(1), Float64(model.coriolis.f)
    elseif model.coriolis !== nothing
        unsupported("Coriolis $(typeof(model.coriolis))")
    end

    bcb, bct = ntuple_fill(OM_BC_DEFAULT, OM_MAX_FIELDS), ntuple_fill(OM_BC_DEFAULT, OM_MAX_FIELDS)
    vb, vt = ntuple_fill(0.0, OM_MAX_FIELDS), ntuple_fill(0.0, OM_MAX_FIELDS)
    fields = (model.velocities.u, model.velocities.v, model.velocities.w, values(model.tracers)...)
    for (q, fld) in enumerate(fields)
        fbcs = fld.boundary_conditions
        kb, xb = bc_code(fbcs.bottom)
        kt, xt = bc_code(fbcs.top)
        for side in (fbcs.west, fbcs.east, fbcs.south, fbcs.north)
            side === nothing || side.classification isa Oceananigans.BoundaryConditions.Periodic ||
                unsupported("boundary conditions on x or y")
        end
        bcb, vb = ntuple_set(bcb, q, kb), ntuple_set(vb, q, xb)
        bct, vt = ntuple_set(bct, q, kt), ntuple_set(vt, q, xt)
    end

    solver = model.pressure_solver
    ps, rt, at, mi, pc = OM_SOLVER_FFT, -1.0, -1.0, Int32(0), OM_PRECOND_DEFAULT
    if solver isa MI355XPoissonSolver
        ps = solver.kind
    elseif solver isa FFTBasedPoissonSolver
        ps = OM_SOLVER_FFT
    elseif solver isa FourierTridiagonalPoissonSolver
        ps = OM_SOLVER_FOURIER_TRIDIAGONAL
    elseif solver isa ConjugateGradientPoissonSolver
        cg = solver.conjugate_gradient_solver
        ps, rt, at, mi = OM_SOLVER_CONJUGATE_GRADIENT, Float64(cg.reltol), Float64(cg.abstol), Int32(min(cg.maxiter, typemax(Int32)))
        pc = preconditioner_code(cg.preconditioner)
    else
        unsupported("pressure solver $(typeof(solver))")
    end

    cfg = OmConfig(FT == Float32 ? Int32(4) : Int32(8), N, H, topo, origin, extent, zfaces,
                   advection_code(model.advection), cl, cp, b, bt, Tt, St, g, α, β, co, f, Int32(nt),
                   bcb, bct, vb, vt, ps, model.pressures.pHY′ === nothing ? Int32(0) : Int32(1),
                   rt, at, mi, pc, bottom, ntuple_fill(Int32(0), 8))
    return cfg, keep
end

# --- model binding -----------------------------------------------------------------------------
# The handle lives beside the Julia model; a finalizer destroys it.
mutable struct MI355XHandle
    ptr::Ptr{Cvoid}
    function MI355XHandle(model)
        cfg, keep = om_config(model)
        h = GC.@preserve keep new(om_create(cfg))
        finalizer(x -> om_destroy(x.ptr), h)
        return h
    end
end

const HANDLES = IdDict{Any,MI355XHandle}()

handle(model) = get!(() -> MI355XHandle(model), HANDLES, model)

const MI355XModel = Oceananigans.AbstractModel{<:Any, MI355X}

prognostic_fields(model) = (model.velocities.u, model.velocities.v, model.velocities.w, values(model.tracers)...)
field_ids(model) = Int32.(0:(2 + length(model.tracers)))

# host Field interior → device (halos are filled by the library)
function push_state!(model)
    h = handle(model)
    for (id, fld) in zip(field_ids(model), prognostic_fields(model))
        om_set_field!(h.ptr, id, Array{Float64,3}(interior(fld)), Int32(0))
    end
    return nothing
end

"""
    pull!(model)

Copy the device state (prognostic fields, pNHS, pHY′) and the clock into the model's host mirrors.
"""
function pull!(model)
    h = handle(model)
    for (id, fld) in zip(field_ids(model), prognostic_fields(model))
        a = Array{Float64,3}(undef, size(interior(fld)))
        om_get_field!(a, h.ptr, id, Int32(0))
        interior(fld) .= a
    end
    for (id, p) in ((OM_FIELD_PNHS, model.pressures.pNHS), (OM_FIELD_PHY, model.pressures.pHY′))
        p === nothing && continue
        a = Array{Float64,3}(undef, size(interior(p)))
        om_get_field!(a, h.ptr, id, Int32(0))
        interior(p) .= a
    end
    pull_clock!(model, h)
    return nothing
end

# Sync the Julia clock after a library step (clock.jl:147-173).
function pull_clock!(model, h)
    c = om_get_clock(h.ptr)
    model.clock.time = c.time
    model.clock.iteration = c.iteration
    model.clock.stage = c.stage
    model.clock.last_Δt = c.last_Δt
    model.clock.last_stage_Δt = c.last_stage_Δt
    return nothing
end

# time_step! for models on the MI355X architecture: the whole RK3 step runs in the library.
function time_step!(model::Oceananigans.AbstractModel{<:RungeKutta3TimeStepper, MI355X}, Δt; callbacks = [])
    isempty(callbacks) || throw(ArgumentError("MI355X: per-substage callbacks are not supported"))
    h = handle(model)
    om_time_step!(h.ptr, Δt, 1)
    model.architecture.sync ? pull!(model) : pull_clock!(model, h)
    return nothing
end

# update_state! (called by the constructor, nonhydrostatic_model.jl:309): the device state is the
# model's; the host mirrors are pushed the first time, when the handle is created.
function update_state!(model::MI355XModel, callbacks = []; compute_tendencies = true)
    created = !haskey(HANDLES, model)
    h = handle(model)
    created && push_state!(model)
    om_update_state!(h.ptr)
    return nothing
end

# set!(model; kw...) (set_nonhydrostatic_model.jl:39-74): set the host Fields with the reference's
# own set!, push them, then the library's epilogue (halos, update_state!, projection with Δt = 1,
# update_state!) and pull the projected state back.
function Oceananigans.set!(model::Oceananigans.Models.NonhydrostaticModels.NonhydrostaticModel{<:Any, <:Any, MI355X};
                           enforce_incompressibility = true, kwargs...)
    for (name, value) in kwargs
        fld = name in keys(model.velocities) ? getproperty(model.velocities, name) :
              name in keys(model.tracers) ? getproperty(model.tracers, name) :
              throw(ArgumentError("name $name not found in model.velocities or model.tracers."))
        Oceananigans.set!(fld, value)
    end
    h = handle(model)
    push_state!(model)
    om_fill_halos!(h.ptr, typemax(UInt32), Int32(1))
    om_update_state!(h.ptr)
    if enforce_incompressibility
        om_pressure_correction!(h.ptr, 1.0)
        om_update_state!(h.ptr)
    end
    pull!(model)
    return nothing
end

end # module

# OceananigansMI355XExt.jl — Julia package extension that routes the NonhydrostaticModel RK3 time
# step on a RectilinearGrid to liboceananigans_mi355x.so through its C ABI.
#
# Seams used (paths relative to the Oceananigans.jl tree):
#   architecture      src/Architectures.jl:21-54 (AbstractSerialArchitecture)
#   time stepping     src/TimeSteppers/runge_kutta_3.jl:103-168 (time_step!), :224 (rk3_substep!)
#   state update      src/Models/NonhydrostaticModels/update_nonhydrostatic_model_state.jl:22-62
#   set!              src/Models/NonhydrostaticModels/set_nonhydrostatic_model.jl:39-74
# The Reactant extension overrides the same seam (ext/OceananigansReactantExt/TimeSteppers.jl:22-30).
module OceananigansMI355XExt

using Oceananigans
using Oceananigans.Architectures: AbstractSerialArchitecture
using Oceananigans.Grids: RectilinearGrid, Periodic, Bounded, Flat, topology, halo_size, znodes, Face
using Oceananigans.TimeSteppers: RungeKutta3TimeStepper
import Oceananigans.TimeSteppers: time_step!, update_state!

const libom = get(ENV, "OCEANANIGANS_MI355X_LIB", "liboceananigans_mi355x")

const OM_MAX_FIELDS = 11

struct MI355X <: AbstractSerialArchitecture end

# Mirrors om_config_t in include/oceananigans_mi355x.h (layout checked by tests/test_abi.py on the
# C side; Julia's isbits struct layout follows the same C alignment rules).
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
om_check(rc, h) = rc == 0 || throw(ErrorException(unsafe_string(ccall((:om_last_error, libom), Cstring, (Ptr{Cvoid},), h))))

function om_create(cfg::OmConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:om_create, libom), Cint, (Ref{OmConfig}, Ref{Ptr{Cvoid}}), cfg, h)
    rc == 0 || throw(ArgumentError(unsafe_string(ccall((:om_last_error, libom), Cstring, (Ptr{Cvoid},), C_NULL))))
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

# --- model binding -----------------------------------------------------------------------------
# The handle lives beside the Julia model; a finalizer destroys it.
mutable struct MI355XHandle
    ptr::Ptr{Cvoid}
    function MI355XHandle(cfg::OmConfig)
        h = new(om_create(cfg))
        finalizer(x -> om_destroy(x.ptr), h)
        return h
    end
end

const HANDLES = IdDict{Any,MI355XHandle}()

handle(model) = get!(() -> MI355XHandle(om_config(model)), HANDLES, model)

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
    pull_clock!(model, h)
    return nothing
end

update_state!(model::Oceananigans.AbstractModel{<:Any, MI355X}; callbacks = []) = om_update_state!(handle(model).ptr)

# om_config(model) translates the model's grid, advection, closure, buoyancy, Coriolis, tracers and
# boundary conditions into an OmConfig, throwing ArgumentError for anything outside the supported
# set (forcing, Stokes drift, particles, non-RectilinearGrid, Bounded x/y).  It is the Julia twin
# of build_config in oceananigans_mi355x/models.py.
function om_config end

end # module

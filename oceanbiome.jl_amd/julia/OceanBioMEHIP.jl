"""
    OceanBioMEHIP

Julia side of libobme_hip.so: a `HIPBiogeochemistry <: AbstractBiogeochemistry`
that Oceananigans drives through the `Oceananigans.Biogeochemistry` hooks
(src/OceanBioME.jl:122-169), with the light scan, tendencies, sinking and CO₂
flux computed by the MI355X kernels through `ccall`.

  - `required_biogeochemical_tracers` and `required_biogeochemical_auxiliary_fields`
    return the wrapped model's names (NutrientsPlanktonDetritus.jl:69-74,
    Light/2band.jl:160);
  - `update_biogeochemical_state!` calls `obme_update_state` (PAR scan);
  - `update_tendencies!` calls `obme_add_tendencies`, which ADDS the
    biogeochemical transitions, minus the drift flux divergence and the top
    CO₂ flux, into `model.timestepper.Gⁿ` (the pattern of
    src/Sediments/tracer_coupling.jl:3-38);
  - the per-cell functor returns zero and `biogeochemical_drift_velocity`
    returns `nothing`, so Oceananigans adds nothing twice.

Field data passed to the library is `parent(field)` (Oceananigans layout,
halos included) of a ROCArray; the pointers are device pointers.
"""
module OceanBioMEHIP

using Oceananigans
using Oceananigans.Biogeochemistry: AbstractBiogeochemistry
import Oceananigans.Biogeochemistry: required_biogeochemical_tracers,
                                     required_biogeochemical_auxiliary_fields,
                                     biogeochemical_auxiliary_fields,
                                     update_biogeochemical_state!,
                                     update_tendencies!,
                                     biogeochemical_drift_velocity

const libobme = get(ENV, "OBME_LIBRARY", joinpath(@__DIR__, "..", "oceanbiome", "libobme_hip.so"))

const OBME_OK = Int32(0)
const OBME_NPZD = Int32(1)
const OBME_LOBSTER = Int32(2)
const OBME_FP64 = Int32(0)
const OBME_FP32_MIXED = Int32(1)
const OBME_MAX_TRACERS = 32

# ---- ABI structs (field order and padding as in include/obme.h) -------------

struct ObmeNPDParams
    values :: NTuple{30, Float64}   # nitrate_half_saturation … small_particle_fraction
    light_limitation :: Int32
    use_temperature_coefficient :: Int32
    phytoplankton_mortality_formulation :: Int32
    grazing_concentration_formulation :: Int32
end

struct ObmeTwoBandPAR
    coefficients :: NTuple{8, Float64}   # kʳ, kᵇ, χʳ, χᵇ, eʳ, eᵇ, r, Rᶜₚ
    surface_mode :: Int32
    pad0 :: Int32
    surface_value :: Float64
    surface_array :: Ptr{Cvoid}
end

struct ObmeSinking
    speed :: NTuple{OBME_MAX_TRACERS, Float64}
    open_bottom :: Int32
    scheme :: Int32
    smoothing_distance :: Float64
end

struct ObmeCarbonSolver
    max_iters :: Int32
    pad0 :: Int32
    atol :: Float64
    damping :: Float64
    armijo_constant :: Float64
    min_damping :: Float64
    initial_pH_guess :: Float64
end

struct ObmeCO2Exchange
    enabled :: Int32
    pad0 :: Int32
    air_concentration :: Float64
    wind_speed :: Float64
    air_pressure :: Float64
    schmidt :: NTuple{5, Float64}
    transfer_velocity :: NTuple{3, Float64}
end

# OxygenGasExchangeBoundaryCondition (GasExchange/GasExchange.jl:127-146)
struct ObmeO2Exchange
    enabled :: Int32
    pad0 :: Int32
    air_concentration :: Float64     # 9352.7 mmol O₂ / m³
    wind_speed :: Float64
    schmidt :: NTuple{5, Float64}
    transfer_velocity :: NTuple{3, Float64}
    solubility :: NTuple{6, Float64}
end

struct ObmeGrid
    Nx :: Int32; Ny :: Int32; Nz :: Int32
    Hx :: Int32; Hy :: Int32; Hz :: Int32
    zc :: Ptr{Float64}
    zf :: Ptr{Float64}
    dz :: Ptr{Float64}
end

struct ObmeFieldSource
    mode :: Int32
    pad0 :: Int32
    value :: Float64
    profile :: Ptr{Float64}
end

const OBME_MAX_SCALE_GROUPS = 5
const OBME_MAX_SCALE_MEMBERS = 12

# ScaleNegativeTracers groups (src/Utils/negative_tracers.jl); n_groups = 0: off.
# `member` and `scalefactor` flatten the C arrays [group][member]: entry
# g * OBME_MAX_SCALE_MEMBERS + m + 1 (0-based g, m); members are 0-based
# tracer indices in required_biogeochemical_tracers order.
struct ObmeNegativeScaling
    n_groups :: Int32
    n_members :: NTuple{OBME_MAX_SCALE_GROUPS, Int32}
    member :: NTuple{OBME_MAX_SCALE_GROUPS * OBME_MAX_SCALE_MEMBERS, Int32}
    pad0 :: Int32
    scalefactor :: NTuple{OBME_MAX_SCALE_GROUPS * OBME_MAX_SCALE_MEMBERS, Float64}
    invalid_fill_value :: NTuple{OBME_MAX_SCALE_GROUPS, Float64}
end

const OBME_MAX_SEDIMENT_TRACERS = 6
const OBME_SEDIMENT_INSTANT_REMINERALISATION = Int32(1)
const OBME_SEDIMENT_SIMPLE_MULTI_G = Int32(2)

# InstantRemineralisationSediment (Models/Sediments/instant_remineralisation.jl) or
# SimpleMultiGSediment's nitrogen form (Models/Sediments/simple_multi_G.jl): tracers as
# 0-based indices in required_biogeochemical_tracers order; model = 0: no sediment.
struct ObmeSediment
    model :: Int32
    receiver :: Int32
    n_sinking :: Int32
    sinking :: NTuple{OBME_MAX_SEDIMENT_TRACERS, Int32}
    timestepper :: Int32             # 0: QuasiAdamsBashforth2
    burial_efficiency_constant1 :: Float64
    burial_efficiency_constant2 :: Float64
    burial_efficiency_half_saturation :: Float64
    ab2_chi :: Float64
    ammonia :: Int32                 # SimpleMultiG: NH₄, NO₃, O₂
    nitrate :: Int32
    oxygen :: Int32
    fast_decay_rate :: Float64
    slow_decay_rate :: Float64
    fast_fraction :: Float64
    slow_fraction :: Float64
    refactory_fraction :: Float64
    sedimentation_rate :: Float64
    anoxia_half_saturation :: Float64
    sinking_redfield :: Float64
    nitrate_oxidation_params :: NTuple{6, Float64}
    denitrification_params :: NTuple{6, Float64}
    anoxic_params :: NTuple{6, Float64}
end

# Oxygen() (NutrientsPlanktonDetritus/oxygen.jl): the O₂ tracer of NPZD or LOBSTER, last in
# the tracer order; enabled = 0: no O₂.
struct ObmeOxygen
    enabled :: Int32
    pad0 :: Int32
    respiration_oxygen_nitrogen_ratio :: Float64     # 10.75
    nitrification_oxygen_nitrogen_ratio :: Float64   # 2.0
end

struct ObmeConfig
    model :: Int32
    precision :: Int32
    carbonate_system :: Int32
    pad0 :: Int32
    grid :: ObmeGrid
    npd :: ObmeNPDParams
    par :: ObmeTwoBandPAR
    sinking :: ObmeSinking
    solver :: ObmeCarbonSolver
    co2 :: ObmeCO2Exchange
    temperature :: ObmeFieldSource
    salinity :: ObmeFieldSource
    scaling :: ObmeNegativeScaling
    sediment :: ObmeSediment
    oxygen :: ObmeOxygen
    o2 :: ObmeO2Exchange             # the O₂ gas exchange of LOBSTER's O₂
end

# layout check against the library (obme_abi_sizeof)
function check_abi()
    for (idx, T) in enumerate((ObmeConfig, ObmeGrid, ObmeNPDParams, ObmeTwoBandPAR, ObmeSinking,
                               ObmeCarbonSolver, ObmeCO2Exchange, ObmeFieldSource))
        n = ccall((:obme_abi_sizeof, libobme), Int64, (Int32,), idx - 1)
        n == sizeof(T) || error("ABI mismatch for $T: library $n bytes, Julia $(sizeof(T))")
    end
    for (idx, T) in ((10, ObmeNegativeScaling), (11, ObmeO2Exchange), (12, ObmeSediment), (13, ObmeOxygen))
        n = ccall((:obme_abi_sizeof, libobme), Int64, (Int32,), idx)
        n == sizeof(T) || error("ABI mismatch for $T: library $n bytes, Julia $(sizeof(T))")
    end
    return nothing
end

last_error() = unsafe_string(ccall((:obme_last_error, libobme), Cstring, ()))
check(status) = status == OBME_OK || error("libobme_hip: status $status: $(last_error())")

# ---- the wrapped biogeochemistry ----------------------------------------------

mutable struct HIPBiogeochemistry{B, F} <: AbstractBiogeochemistry
    underlying :: B               # the OceanBioME model it replaces (names, parameters)
    config :: Base.RefValue{ObmeConfig}
    handle :: Ptr{Cvoid}
    PAR :: F                      # CenterField owned by the wrapper
    keep :: Vector{Any}           # host arrays referenced by `config`
    storage :: Any                # the sediment's fields (device Float64) or nothing
end

"""
    HIPBiogeochemistry(underlying, grid, config::ObmeConfig, PAR; keep = [], storage = nothing)

Creates the library handle for `config` (built from the OceanBioME model
`underlying` by the caller, see the Python host package for the field-by-field
mapping) and owns `PAR`.  With a sediment in `config`, `storage` holds the
sediment's fields as a device array of Float64, i fastest:
InstantRemineralisation's storage (Nx*Ny), or SimpleMultiG's Ns, Nf, Nr one
after the other (3 Nx*Ny).
"""
function HIPBiogeochemistry(underlying, grid, config::ObmeConfig, PAR; keep = Any[], storage = nothing)
    check_abi()
    h = Ref{Ptr{Cvoid}}(C_NULL)
    cfg = Ref(config)
    check(ccall((:obme_create, libobme), Int32, (Ref{ObmeConfig}, Ref{Ptr{Cvoid}}), cfg, h))
    bgc = HIPBiogeochemistry(underlying, cfg, h[], PAR, keep, storage)
    finalizer(b -> ccall((:obme_destroy, libobme), Int32, (Ptr{Cvoid},), b.handle), bgc)
    return bgc
end

required_biogeochemical_tracers(b::HIPBiogeochemistry) = required_biogeochemical_tracers(b.underlying)
required_biogeochemical_auxiliary_fields(::HIPBiogeochemistry) = (:PAR, )
biogeochemical_auxiliary_fields(b::HIPBiogeochemistry) = (; PAR = b.PAR)
biogeochemical_drift_velocity(::HIPBiogeochemistry, ::Val) = nothing

# the per-cell functor: the library adds the transitions in update_tendencies!
@inline (::HIPBiogeochemistry)(i, j, k, grid, ::Val, clock, fields) = zero(eltype(grid))

device_pointer(field) = Ptr{Cvoid}(UInt(pointer(parent(field))))

function tracer_pointers(b::HIPBiogeochemistry, tracers)
    names = required_biogeochemical_tracers(b)
    return Ptr{Cvoid}[device_pointer(tracers[name]) for name in names]
end

stream_pointer() = C_NULL   # the default stream of the task (AMDGPU's current stream may be passed instead)

function update_biogeochemical_state!(b::HIPBiogeochemistry, model)
    ptrs = tracer_pointers(b, model.tracers)
    check(ccall((:obme_update_state, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Float64, Ptr{Cvoid}),
                b.handle, ptrs, device_pointer(b.PAR), model.clock.time, stream_pointer()))
    if b.config[].sediment.model != 0
        # update_biogeochemical_state!(model, sediment) (Sediments/update_state.jl:6-17)
        check(ccall((:obme_sediment_update_state, libobme), Int32,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Float64, Ptr{Cvoid}),
                    b.handle, ptrs, Ptr{Cvoid}(UInt(pointer(b.storage))), model.clock.last_stage_Δt,
                    stream_pointer()))
    end
    return nothing
end

function update_tendencies!(b::HIPBiogeochemistry, model)
    ptrs = tracer_pointers(b, model.tracers)
    G = tracer_pointers(b, model.timestepper.Gⁿ)
    T = haskey(model.tracers, :T) && !(:T in required_biogeochemical_tracers(b)) ?
        device_pointer(model.tracers.T) : C_NULL
    S = haskey(model.tracers, :S) ? device_pointer(model.tracers.S) : C_NULL
    if b.config[].sediment.model == OBME_SEDIMENT_SIMPLE_MULTI_G
        # the coupled NH₄, NO₃, O₂ of SimpleMultiG from its fields (tracer_coupling.jl:3-38)
        check(ccall((:obme_add_tendencies_with_sediment, libobme), Int32,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                     Float64, Ptr{Cvoid}),
                    b.handle, ptrs, G, device_pointer(b.PAR), T, S, Ptr{Cvoid}(UInt(pointer(b.storage))),
                    model.clock.time, stream_pointer()))
        return nothing
    end
    check(ccall((:obme_add_tendencies, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}),
                b.handle, ptrs, G, device_pointer(b.PAR), T, S, model.clock.time, stream_pointer()))
    return nothing
end

"""
    step!(b::HIPBiogeochemistry, tracers, nsteps, Δt; t₀ = 0)

The library's fused RK3 stepper with frozen physics (update-state,
tendencies and substep per stage), in place on the tracer fields.
"""
function step!(b::HIPBiogeochemistry, tracers, nsteps, Δt; t₀ = 0.0)
    ptrs = tracer_pointers(b, tracers)
    if b.config[].sediment.model != 0
        check(ccall((:obme_step_with_sediment, libobme), Int32,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Float64, Float64,
                     Ptr{Cvoid}),
                    b.handle, ptrs, C_NULL, C_NULL, Ptr{Cvoid}(UInt(pointer(b.storage))), Int32(nsteps),
                    Float64(Δt), Float64(t₀), stream_pointer()))
        return nothing
    end
    check(ccall((:obme_step, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Float64, Float64, Ptr{Cvoid}),
                b.handle, ptrs, C_NULL, C_NULL, Int32(nsteps), Float64(Δt), Float64(t₀), stream_pointer()))
    return nothing
end

# ---- PISCES ---------------------------------------------------------------------

const OBME_PP_COUNT = 159

# OxygenGasExchangeBoundaryCondition (GasExchange.jl:127-146)
struct ObmeO2Exchange
    enabled :: Int32
    pad0 :: Int32
    air_concentration :: Float64
    wind_speed :: Float64
    schmidt :: NTuple{5, Float64}
    transfer_velocity :: NTuple{3, Float64}
    solubility :: NTuple{6, Float64}
end

struct ObmePiscesConfig
    precision :: Int32
    open_bottom :: Int32
    grid :: ObmeGrid
    params :: NTuple{OBME_PP_COUNT, Float64}      # OBME_PP_* layout of include/obme.h
    band_kw :: NTuple{3, Float64}
    band_e :: NTuple{3, Float64}
    band_chi :: NTuple{3, Float64}
    band_division :: NTuple{3, Float64}
    surface_PAR :: Float64
    surface_mode :: Int32            # OBME_SURFACE_CONSTANT | _DIURNAL | _ARRAY
    surface_array :: Ptr{Cvoid}      # OBME_SURFACE_ARRAY: Nx×Ny device values
    solver :: ObmeCarbonSolver
    scaling :: ObmeNegativeScaling
    co2 :: ObmeCO2Exchange           # top DIC flux (validation/PISCES/column.jl:63)
    o2 :: ObmeO2Exchange             # top O₂ flux (:64)
    sediment :: ObmeSediment         # PISCES(; sediment): InstantRemineralisation or none
end

"""
    HIPPISCES(underlying, config::ObmePiscesConfig, aux)

`PISCES(; grid)` (PISCES.jl:268-330) on the device.  `aux` holds the
auxiliary fields the library writes in `update_biogeochemical_state!`:
`PAR`, `PAR₁`, `PAR₂`, `PAR₃`, `Ω` (CenterFields), `zₑᵤ` and
`mixed_layer_PAR` (2-D, Nx×Ny).  With a sediment in `config` (PISCES's
InstantRemineralisation), `storage` is its Nx×Ny device array of Float64.
"""
mutable struct HIPPISCES{B, A} <: AbstractBiogeochemistry
    underlying :: B
    config :: Base.RefValue{ObmePiscesConfig}
    handle :: Ptr{Cvoid}
    aux :: A
    storage :: Any                   # the sediment's storage (device Float64) or nothing
end

function HIPPISCES(underlying, config::ObmePiscesConfig, aux; storage = nothing)
    n = ccall((:obme_abi_sizeof, libobme), Int64, (Int32,), 9)
    n == sizeof(ObmePiscesConfig) || error("ABI mismatch for ObmePiscesConfig")
    n = ccall((:obme_abi_sizeof, libobme), Int64, (Int32,), 11)
    n == sizeof(ObmeO2Exchange) || error("ABI mismatch for ObmeO2Exchange")
    h = Ref{Ptr{Cvoid}}(C_NULL)
    cfg = Ref(config)
    check(ccall((:obme_pisces_create, libobme), Int32, (Ref{ObmePiscesConfig}, Ref{Ptr{Cvoid}}), cfg, h))
    bgc = HIPPISCES(underlying, cfg, h[], aux, storage)
    finalizer(b -> ccall((:obme_pisces_destroy, libobme), Int32, (Ptr{Cvoid},), b.handle), bgc)
    return bgc
end

required_biogeochemical_tracers(b::HIPPISCES) = required_biogeochemical_tracers(b.underlying)
required_biogeochemical_auxiliary_fields(::HIPPISCES) = (:PAR, :PAR₁, :PAR₂, :PAR₃, :Ω, :zₑᵤ, :mixed_layer_PAR)
biogeochemical_auxiliary_fields(b::HIPPISCES) = b.aux
biogeochemical_drift_velocity(::HIPPISCES, ::Val) = nothing
@inline (::HIPPISCES)(i, j, k, grid, ::Val, clock, fields) = zero(eltype(grid))

function update_biogeochemical_state!(b::HIPPISCES, model)
    ptrs = Ptr{Cvoid}[device_pointer(model.tracers[name]) for name in required_biogeochemical_tracers(b)]
    bands = Ptr{Cvoid}[device_pointer(b.aux.PAR₁), device_pointer(b.aux.PAR₂), device_pointer(b.aux.PAR₃)]
    check(ccall((:obme_pisces_update_state, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Float64, Ptr{Cvoid}),
                b.handle, ptrs, device_pointer(b.aux.PAR), bands, device_pointer(b.aux.Ω),
                Ptr{Cvoid}(UInt(pointer(b.aux.zₑᵤ))), Ptr{Cvoid}(UInt(pointer(b.aux.mixed_layer_PAR))),
                model.clock.time, stream_pointer()))
    if b.config[].sediment.model != 0
        # update_biogeochemical_state!(model, sediment) (Sediments/update_state.jl:6-17)
        check(ccall((:obme_pisces_sediment_update_state, libobme), Int32,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Float64, Ptr{Cvoid}),
                    b.handle, ptrs, Ptr{Cvoid}(UInt(pointer(b.storage))), model.clock.last_stage_Δt,
                    stream_pointer()))
    end
    return nothing
end

function update_tendencies!(b::HIPPISCES, model)
    names = required_biogeochemical_tracers(b)
    ptrs = Ptr{Cvoid}[device_pointer(model.tracers[name]) for name in names]
    G = Ptr{Cvoid}[device_pointer(model.timestepper.Gⁿ[name]) for name in names[1:24]]
    check(ccall((:obme_pisces_add_tendencies, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Float64, Ptr{Cvoid}),
                b.handle, ptrs, G, model.clock.time, stream_pointer()))
    return nothing
end

function step!(b::HIPPISCES, tracers, nsteps, Δt; t₀ = 0.0)
    ptrs = Ptr{Cvoid}[device_pointer(tracers[name]) for name in required_biogeochemical_tracers(b)]
    if b.config[].sediment.model != 0
        check(ccall((:obme_pisces_step_with_sediment, libobme), Int32,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Int32, Float64, Float64, Ptr{Cvoid}),
                    b.handle, ptrs, Ptr{Cvoid}(UInt(pointer(b.storage))), Int32(nsteps), Float64(Δt),
                    Float64(t₀), stream_pointer()))
        return nothing
    end
    check(ccall((:obme_pisces_step, libobme), Int32,
                (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Int32, Float64, Float64, Ptr{Cvoid}),
                b.handle, ptrs, Int32(nsteps), Float64(Δt), Float64(t₀), stream_pointer()))
    return nothing
end

export HIPBiogeochemistry, HIPPISCES, step!

end # module

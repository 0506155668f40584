"""
    FastIceHIP

FastIce.jl's pseudo-transient Stokes path on an AMD MI355X through
libfasticehip.so (C ABI: include/fastice_hip.h).  The functions keep the
reference's names and signatures, so a driver such as
scripts_future_API/test_vavilov.jl switches by changing its `using` line:

  * `IsothermalFullStokesModel(; arch, grid, boundary_conditions, gravity,
    rheology, solver_params, outer_width)`
    (src/Models/FullStokes/Isothermal/Isothermal.jl:60-89) and
    `IsothermalImmersedBoundaryFullStokesModel(; ..., level_sets)`
    (src/Models/ImmersedBoundaryFullStokes/Isothermal/Isothermal.jl:63-98);
  * `advance_iteration!(model, t, Δt)` (Isothermal.jl:91-112),
    `compute_residuals!(model)` (Isothermal.jl:114-125);
  * `compute_levelset_from_dem!(arch, launch, Ψ, dem, dem_grid2D, grid, R)`
    and `invert_levelset!(arch, launch, Ψ, grid)`
    (src/LevelSets/compute_levelsets.jl:21-35);
  * `ω_from_ψ!(arch, launch, ω, ψ, grid)` (volume_fractions_kernels.jl:101-110);
  * `set!`, `interior`, `bc!(arch, grid, f => Neumann())` and `norm(f, Inf)`
    on the model's fields (scripts_future_API/tm_stokes_chmy.jl:86-113).

The fields live on the device, owned by the library; `model.stress.P`,
`model.velocity.V.x`, `model.viscosity.η`, `model.residual.r_P`,
`model.level_sets.ns`, `model.field_masks.na`, ... are handles.  Rheology
and gravity callables cannot cross a C ABI: `lower_desc` lowers them to the
fi_rheology / fi_model_desc records (constant or field parameters, the
viscoplastic cap and the relaxation scale).
"""
module FastIceHIP

using LinearAlgebra
using FastIce.Physics: LinearViscousRheology, PowerLawRheology, ViscoplasticRegularisation, Relaxation
using FastIce.Models.FullStokes.Isothermal: BoundaryCondition, Traction, Velocity, Slip
using Chmy.Grids: Center, Vertex, origin, extent

export IsothermalFullStokesModel, IsothermalImmersedBoundaryFullStokesModel, advance_iteration!, compute_residuals!
export compute_levelset_from_dem!, invert_levelset!, ω_from_ψ!, set!, interior, bc!, Neumann, lower_desc

const lib = get(ENV, "FASTICE_HIP_LIB", joinpath(@__DIR__, "..", "..", "..", "fastice.jl_amd", "libfasticehip.so"))

# --- records of include/fastice_hip.h (field order and types as declared) ----

const FI_BC_TRACTION, FI_BC_VELOCITY, FI_BC_SLIP = Int32(0), Int32(1), Int32(2)
const FI_RHEO_LINEAR_CONST, FI_RHEO_LINEAR_FIELD, FI_RHEO_POWERLAW = Int32(0), Int32(1), Int32(2)
const FI_RELAX_NONE, FI_RELAX_LIN, FI_RELAX_LOG, FI_RELAX_LOG10 = Int32(0), Int32(1), Int32(2), Int32(3)
const FI_LAYOUT_PARENT, FI_LAYOUT_INTERIOR = Int32(0), Int32(1)
const FI_BCGROUP_STRESS, FI_BCGROUP_VELOCITY, FI_BCGROUP_ETA = Int32(0), Int32(1), Int32(2)
const FI_PHASE_NA, FI_PHASE_NS = Int32(0), Int32(1)
const FIELD_IDS = (P=0, TXX=1, TYY=2, TZZ=3, TXY=4, TXZ=5, TYZ=6, VX=7, VY=8, VZ=9, ETA=10, ETA_NEXT=11,
                   RP=12, RVX=13, RVY=14, RVZ=15, GX=16, GY=17, GZ=18, RHEO=19, PSI_NA=20, PSI_NS=21,
                   OM_NA0=22, OM_NS0=30)

struct FiRheology
    kind::Int32
    has_reg::Int32
    relax::Int32
    A_is_field::Int32
    eta::Float64
    A::Float64
    n::Float64
    eta_reg::Float64
    rate::Float64
end

struct FiModelDesc
    ndims::Int32
    n::NTuple{3,Int32}
    origin::NTuple{3,Float64}
    extent::NTuple{3,Float64}
    bc_kind::NTuple{3,NTuple{2,Int32}}
    bc_value::NTuple{3,NTuple{2,NTuple{3,Float64}}}
    rheology::FiRheology
    gravity_is_field::Int32
    gravity::NTuple{3,Float64}
    dtau_P::Float64
    dtau_tau::NTuple{6,Float64}
    dtau_V::NTuple{3,Float64}
    immersed_boundary::Int32
    precision::Int32
end

struct FiError <: Exception
    code::Int32
    msg::String
end
Base.showerror(io::IO, e::FiError) = print(io, "libfasticehip error ", e.code, ": ", e.msg)

function check(rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:fi_last_error, lib), Cstring, ()))
    # tm_stokes_chmy.jl:104-106: a non-finite residual is "simulation failed"
    rc == -4 && error("simulation failed: ", msg)
    throw(FiError(rc, msg))
end

# --- lowering of the reference's model arguments ---------------------------

bc_code(::BoundaryCondition{Traction}) = FI_BC_TRACTION
bc_code(::BoundaryCondition{Velocity}) = FI_BC_VELOCITY
bc_code(::BoundaryCondition{Slip}) = FI_BC_SLIP

# a parameter that is a constant (a number, a ValueField, ZeroField, OneField
# or ConstantField) or a field sampled at the cell centres / faces
function constant_value(p)
    p isa Number && return Float64(p)
    hasproperty(p, :value) && return Float64(getproperty(p, :value))
    name = string(nameof(typeof(p)))
    name == "ZeroField" && return 0.0
    name == "OneField" && return 1.0
    return nothing
end

# (inner law, optional ViscoplasticRegularisation, optional Relaxation); the
# relaxed field of a Relaxation is the model's lagged viscosity η
function lower_rheology(r)
    relax, rate = FI_RELAX_NONE, 0.0
    if r isa Relaxation
        relax = isnothing(r.scale) ? FI_RELAX_LIN : r.scale === log ? FI_RELAX_LOG : r.scale === log10 ? FI_RELAX_LOG10 :
                throw(ArgumentError("unsupported relaxation scale $(r.scale)"))
        rate = Float64(r.rate)
        r = r.target
    end
    has_reg, eta_reg = Int32(0), 0.0
    if r isa ViscoplasticRegularisation
        has_reg, eta_reg = Int32(1), 1.0 / Float64(r.inv_η_reg)
        r = r.viscosity
    end
    field = nothing
    if r isa LinearViscousRheology
        c = constant_value(r.η)
        kind, eta, A, n, A_is_field = isnothing(c) ? FI_RHEO_LINEAR_FIELD : FI_RHEO_LINEAR_CONST,
                                      something(c, 0.0), 0.0, 0.0, Int32(0)
        isnothing(c) && (field = r.η)
    elseif r isa PowerLawRheology
        c = constant_value(r.A)
        kind, eta, A, n, A_is_field = FI_RHEO_POWERLAW, 0.0, something(c, 0.0), Float64(r.n), Int32(isnothing(c))
        isnothing(c) && (field = r.A)
    else
        throw(ArgumentError("rheology $(typeof(r)) has no C ABI lowering"))
    end
    return FiRheology(kind, has_reg, relax, A_is_field, eta, A, n, eta_reg, rate), field
end

"""
    lower_desc(grid, boundary_conditions, gravity, rheology, solver_params, immersed_boundary)

The fi_model_desc record of a model and the parts that travel as fields:
(desc, rheology field or nothing, gravity fields (NamedTuple of arrays or
nothing)).  Grid geometry from the Vertex extent; boundary conditions
`(x=(west, east), y=(south, north)[, z=(bottom, top)])` of
`BoundaryCondition{Traction|Velocity|Slip}` (boundary_conditions.jl:8-13);
`solver_params.Δτ = (Pr, τ=(xx, yy[, zz], xy[, xz, yz]), V=(x, y[, z]))`.
"""
function lower_desc(grid, boundary_conditions, gravity, rheology, solver_params, immersed_boundary::Bool)
    nd = ndims(grid)
    dims = size(grid, Center())
    o, e = origin(grid, Vertex()), extent(grid, Vertex())
    pad3(t, z) = ntuple(q -> q <= length(t) ? t[q] : z, 3)
    sides = values(boundary_conditions)
    kinds = ntuple(d -> d <= nd ? (bc_code(sides[d][1]), bc_code(sides[d][2])) : (Int32(0), Int32(0)), 3)
    vals = ntuple(d -> d <= nd ? ntuple(s -> pad3(Float64.(sides[d][s].components), 0.0), 2) :
                                 ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), 3)
    rh, rfield = lower_rheology(rheology)
    gs = values(gravity)
    gconst = map(constant_value, gs)
    gfield = any(isnothing, gconst)
    gvec = pad3(Tuple(something(c, 0.0) for c in gconst), 0.0)
    Δτ = solver_params.Δτ
    τ = nd == 3 ? (Δτ.τ.xx, Δτ.τ.yy, Δτ.τ.zz, Δτ.τ.xy, Δτ.τ.xz, Δτ.τ.yz) : (Δτ.τ.xx, Δτ.τ.yy, Δτ.τ.xy, 0.0, 0.0, 0.0)
    V = pad3(Tuple(values(Δτ.V)), 0.0)
    desc = FiModelDesc(Int32(nd), pad3(Int32.(dims), Int32(0)), pad3(Float64.(o), 0.0), pad3(Float64.(e), 1.0),
                       kinds, vals, rh, Int32(gfield), gvec, Float64(Δτ.Pr), Float64.(τ), Float64.(V),
                       Int32(immersed_boundary), Int32(0))
    return desc, rfield, gfield ? gs : nothing
end

# --- models and device fields -----------------------------------------------

mutable struct HIPModel
    h::Ptr{Cvoid}
    grid::Any
    arch::Any
    immersed::Bool
    stress::Any
    velocity::Any
    viscosity::Any
    residual::Any
    level_sets::Any
    field_masks::Any
    boundary_conditions::Any
end

struct HIPField
    model::HIPModel
    id::Int32
    loc::NTuple{3,Int} # 1 = Vertex in that dimension
end

struct HIPMask  # a FieldMask of one phase: its eight staggered components
    model::HIPModel
    phase::Int32
end

struct Neumann end

function field_loc(nd, name)
    name in (:P, :TXX, :TYY, :TZZ, :ETA, :ETA_NEXT, :RP, :RHEO) && return (0, 0, 0)
    name in (:VX, :RVX, :GX) && return (1, 0, 0)
    name in (:VY, :RVY, :GY) && return (0, 1, 0)
    name in (:VZ, :RVZ, :GZ) && return (0, 0, 1)
    name === :TXY && return (1, 1, 0)
    name === :TXZ && return (1, 0, 1)
    name === :TYZ && return (0, 1, 1)
    return nd == 3 ? (1, 1, 1) : (1, 1, 0)  # level sets at vertices
end

hf(m, name) = HIPField(m, Int32(FIELD_IDS[name]), field_loc(ndims(m.grid), name))

interior_size(f::HIPField) = ntuple(d -> size(f.model.grid, Center())[d] + f.loc[d], ndims(f.model.grid))
parent_size(f::HIPField) = ntuple(d -> size(f.model.grid, Center())[d] + 3, ndims(f.model.grid))

"""set!(f, A): an array of the interior or parent size (Julia order), or a number."""
function set!(f::HIPField, v::Number)
    check(ccall((:fi_field_fill, lib), Cint, (Ptr{Cvoid}, Cint, Cdouble), f.model.h, f.id, Float64(v)))
end

function set!(f::HIPField, A::AbstractArray)
    B = Array{Float64}(A)
    layout = size(B) == parent_size(f) ? FI_LAYOUT_PARENT : size(B) == interior_size(f) ? FI_LAYOUT_INTERIOR :
             throw(DimensionMismatch("array of size $(size(B)) fits neither the parent nor the interior"))
    check(ccall((:fi_field_upload, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Csize_t, Cint),
                f.model.h, f.id, B, length(B), layout))
end

"""interior(f): the field's interior as an Array (a host copy)."""
function interior(f::HIPField)
    A = Array{Float64}(undef, interior_size(f))
    check(ccall((:fi_field_download, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Csize_t, Cint),
                f.model.h, f.id, A, length(A), FI_LAYOUT_INTERIOR))
    return A
end

function Base.parent(f::HIPField)
    A = Array{Float64}(undef, parent_size(f))
    check(ccall((:fi_field_download, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Csize_t, Cint),
                f.model.h, f.id, A, length(A), FI_LAYOUT_PARENT))
    return A
end

Base.Array(f::HIPField) = interior(f)
LinearAlgebra.norm(f::HIPField, p::Real=2) = norm(interior(f), p)

# bc!(arch, grid, η => Neumann()) for the viscosity (tm_stokes_chmy.jl:93-94)
function bc!(arch, grid, p::Pair{HIPField,Neumann})
    f = p.first
    f.id in (FIELD_IDS.ETA, FIELD_IDS.ETA_NEXT) ||
        throw(ArgumentError("Neumann conditions apply to the viscosity fields"))
    check(ccall((:fi_apply_bc, lib), Cint, (Ptr{Cvoid}, Cint), f.model.h, FI_BCGROUP_ETA))
end

# the model's own boundary conditions on its stress or velocity
apply_bc!(m::HIPModel, group::Symbol) =
    check(ccall((:fi_apply_bc, lib), Cint, (Ptr{Cvoid}, Cint), m.h,
                group === :stress ? FI_BCGROUP_STRESS : FI_BCGROUP_VELOCITY))

function sample(g, loc, grid)
    A = g isa AbstractArray ? g : [g[I] for I in CartesianIndices(ntuple(d -> size(grid, Center())[d] + loc[d], ndims(grid)))]
    return Array{Float64}(A)
end

function create(arch, grid, boundary_conditions, gravity, rheology, solver_params, immersed)
    desc, rfield, gfields = lower_desc(grid, boundary_conditions, gravity, rheology, solver_params, immersed)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fi_model_create, lib), Cint, (Ref{FiModelDesc}, Ref{Ptr{Cvoid}}), desc, h))
    m = HIPModel(h[], grid, arch, immersed, nothing, nothing, nothing, nothing, nothing, nothing, boundary_conditions)
    finalizer(m) do x
        x.h == C_NULL || ccall((:fi_model_destroy, lib), Cint, (Ptr{Cvoid},), x.h)
        x.h = C_NULL
    end
    nd = ndims(grid)
    m.stress = (P=hf(m, :P), τ=nd == 3 ? (xx=hf(m, :TXX), yy=hf(m, :TYY), zz=hf(m, :TZZ), xy=hf(m, :TXY),
                                        xz=hf(m, :TXZ), yz=hf(m, :TYZ)) :
                                       (xx=hf(m, :TXX), yy=hf(m, :TYY), xy=hf(m, :TXY)))
    m.velocity = (V=nd == 3 ? (x=hf(m, :VX), y=hf(m, :VY), z=hf(m, :VZ)) : (x=hf(m, :VX), y=hf(m, :VY)),)
    m.viscosity = (η=hf(m, :ETA), η_next=hf(m, :ETA_NEXT))
    m.residual = (r_P=hf(m, :RP), r_V=nd == 3 ? (x=hf(m, :RVX), y=hf(m, :RVY), z=hf(m, :RVZ)) : (x=hf(m, :RVX), y=hf(m, :RVY)))
    if immersed
        m.level_sets = (na=hf(m, :PSI_NA), ns=hf(m, :PSI_NS))
        m.field_masks = (na=HIPMask(m, FI_PHASE_NA), ns=HIPMask(m, FI_PHASE_NS))
    end
    isnothing(rfield) || set!(hf(m, :RHEO), sample(rfield, (0, 0, 0), grid))
    if !isnothing(gfields)
        for (d, g) in enumerate(gfields)
            loc = ntuple(q -> q == d ? 1 : 0, 3)
            set!(hf(m, (:GX, :GY, :GZ)[d]), sample(g, loc, grid))
        end
    end
    return m
end

"""
    IsothermalFullStokesModel(; arch, grid, boundary_conditions, gravity, rheology, solver_params, outer_width=nothing)

Isothermal.jl:60-89 on the device.  `outer_width` is accepted for the
signature; one device has no halo exchange for it to hide.
"""
IsothermalFullStokesModel(; arch, grid, boundary_conditions, gravity, rheology, solver_params, outer_width=nothing) =
    create(arch, grid, boundary_conditions, gravity, rheology, solver_params, false)

"""
    IsothermalImmersedBoundaryFullStokesModel(; arch, grid, boundary_conditions, level_sets=nothing, gravity,
                                              rheology, solver_params=(), outer_width=nothing)

ImmersedBoundaryFullStokes Isothermal.jl:63-98; `level_sets = (na=…, ns=…)`
given as arrays are uploaded into `model.level_sets`.
"""
function IsothermalImmersedBoundaryFullStokesModel(; arch, grid, boundary_conditions, level_sets=nothing, gravity,
                                                   rheology, solver_params=(), outer_width=nothing)
    m = create(arch, grid, boundary_conditions, gravity, rheology, solver_params, true)
    if !isnothing(level_sets)
        for k in (:na, :ns)
            ψ = getproperty(level_sets, k)
            ψ isa AbstractArray && set!(getproperty(m.level_sets, k), ψ)
        end
    end
    return m
end

advance_iteration!(m::HIPModel, t, Δt) = check(ccall((:fi_advance_iterations, lib), Cint, (Ptr{Cvoid}, Cint), m.h, 1))

"""compute_residuals!(model): residuals on the device; returns their Inf-norms (r_P, r_V.x, r_V.y[, r_V.z])."""
function compute_residuals!(m::HIPModel)
    norms = zeros(4)
    check(ccall((:fi_compute_residuals, lib), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), m.h, norms))
    return ndims(m.grid) == 3 ? norms : norms[1:3]
end

function phase_of(Ψ::HIPField)
    Ψ.id == FIELD_IDS.PSI_NA && return FI_PHASE_NA
    Ψ.id == FIELD_IDS.PSI_NS && return FI_PHASE_NS
    throw(ArgumentError("not a level set of an immersed-boundary model"))
end

"""
    compute_levelset_from_dem!(arch, launch, Ψ, dem, dem_grid2D, grid, R=LinearAlgebra.I)

compute_levelsets.jl:21-25: the signed distance of every vertex to the DEM
surface (two triangles per DEM cell, cut-off 4 max Δ), with Neumann ghosts.
`dem` is the DEM field on `dem_grid2D` with its one-cell halo (`parent` of a
Vertex field) or that array; `R` a rotation applied to the vertices.
"""
function compute_levelset_from_dem!(arch, launch, Ψ::HIPField, dem, dem_grid2D, grid, R=LinearAlgebra.I)
    D = Array{Float64}(dem isa AbstractArray ? dem : parent(dem))
    dn = size(dem_grid2D, Center())
    size(D) == (dn[1] + 3, dn[2] + 3) || throw(DimensionMismatch("DEM array must include its halo"))
    o = Float64[origin(dem_grid2D, Vertex())...]
    e = Float64[extent(dem_grid2D, Vertex())...]
    Rm = R isa UniformScaling ? C_NULL : vec(permutedims(Matrix{Float64}(R)))  # row-major
    check(ccall((:fi_levelset_from_dem, lib), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                Ψ.model.h, phase_of(Ψ), D, dn[1], dn[2], o, e, Rm))
end

"""invert_levelset!(arch, launch, Ψ, grid) (compute_levelsets.jl:32-35): Ψ ← -Ψ."""
invert_levelset!(arch, launch, Ψ::HIPField, grid) =
    check(ccall((:fi_invert_levelset, lib), Cint, (Ptr{Cvoid}, Cint), Ψ.model.h, phase_of(Ψ)))

"""
    ω_from_ψ!(arch, launch, ω, ψ, grid)

volume_fractions_kernels.jl:101-110: the fraction of every staggered control
volume where ψ < 0, in FP64, for the mask of the same phase as ψ.
"""
function ω_from_ψ!(arch, launch, ω::HIPMask, ψ::HIPField, grid)
    ω.model === ψ.model && ω.phase == phase_of(ψ) ||
        throw(ArgumentError("ω and ψ must be the mask and level set of one phase of one model"))
    check(ccall((:fi_omega_from_psi, lib), Cint, (Ptr{Cvoid}, Cint), ω.model.h, ω.phase))
end

end # module

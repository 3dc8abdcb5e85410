"""
    PolyFieldsHIP

`ccall` binding of `libpolyfields_hip.so` (include/polyfields_hip.h) for PolyFields.jl.
It mirrors the reference's user-facing API -- `FieldSystem`, `add_species!`,
`add_interaction!`, `fieldinit!`, `scft!`, `AndersonUpdater`, `DomainUpdater`,
`density!`, `residuals!`, `free_energy`, `field_error`, `scfstress` -- with every grid held
on one MI355X.  Grids cross the boundary as `Array{Float64,3}` in Julia's own column-major
(x fastest) order, so `pointer(A)` is passed without a transpose.

Source-only in this repository: it is not executed by the test suite (no Julia toolchain
in the build image); the Python ctypes mirror in `polyfields_hip/` exercises the same ABI.
"""
module PolyFieldsHIP

using PolyFields: UnitCell, Monomer, Multiblock, Point, FloryInteraction, EdwardsInteraction,
                  Compressibility, SCFTOptions, SCFTResults, Canonical, Grand

const LIB = get(ENV, "POLYFIELDS_HIP_LIB", joinpath(@__DIR__, "..", "polyfields_hip", "libpolyfields_hip.so"))

const PF_FP64 = Cint(0)
const PF_FP32 = Cint(1)

struct PFError <: Exception
    code::Int
    msg::String
end

function check(rc::Cint)
    if rc < 0
        msg = unsafe_string(ccall((:pf_last_error, LIB), Cstring, ()))
        throw(PFError(rc, msg))
    end
    return nothing
end

mutable struct HIPSystem
    handle::Ptr{Cvoid}
    dims::NTuple{3,Int}
    cell::UnitCell
    mids::Vector{Int}
    species::Vector{Any}
end

"""
    FieldSystem(dims, cell; monomers, ensemble = Canonical, compress = Compressibility(0.0),
                precision = :fp64, device = 0)

Device-resident counterpart of `PolyFields.FieldSystem` (src/system/system.jl:51-74).
"""
function FieldSystem(dims::NTuple{3,<:Integer}, cell::UnitCell; monomers = Monomer[],
                     ensemble = Canonical, compress = Compressibility(0.0), precision = :fp64,
                     device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    p = Vector{Float64}(cell.params)
    check(ccall((:pf_system_create, LIB), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Cint, Cstring, Ptr{Float64}, Cint, Cint,
                 Float64, Cint, Cint),
                h, dims[1], dims[2], dims[3], cell.dim, String(cell.crystal), p, length(p),
                ensemble == Canonical ? 0 : 1, compress.inv_zeta,
                precision == :fp32 ? PF_FP32 : PF_FP64, device))
    sys = HIPSystem(h[], Tuple(Int.(dims)), cell, Int[], Any[])
    finalizer(s -> ccall((:pf_system_destroy, LIB), Cint, (Ptr{Cvoid},), s.handle), sys)
    for mon in monomers
        add_monomer!(sys, mon)
    end
    return sys
end

function add_monomer!(sys::HIPSystem, mon::Monomer)
    check(ccall((:pf_system_add_monomer, LIB), Cint, (Ptr{Cvoid}, Cint, Float64, Float64),
                sys.handle, mon.id, mon.vol, mon.charge))
    push!(sys.mids, mon.id)
    return nothing
end

function add_species!(sys::HIPSystem, chain::Multiblock, val::Real)
    mids = Cint[m for m in chain.mid_block]
    Nb = Cint[n for n in chain.N_block]
    bb = Float64[b for b in chain.b_block]
    idx = Ref{Cint}(0)
    nsu = Ref{Cint}(0)
    check(ccall((:pf_system_add_multiblock, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Float64}, Cint, Float64, Ref{Cint},
                 Ref{Cint}),
                sys.handle, length(mids), mids, Nb, bb, chain.Ns, val, idx, nsu))
    push!(sys.species, chain)
    return nothing
end

function add_species!(sys::HIPSystem, pt::Point, val::Real)
    idx = Ref{Cint}(0)
    check(ccall((:pf_system_add_point, LIB), Cint, (Ptr{Cvoid}, Cint, Float64, Ref{Cint}),
                sys.handle, pt.mids[1], val, idx))
    push!(sys.species, pt)
    return nothing
end

function add_interaction!(sys::HIPSystem, itx::FloryInteraction)
    for ((a, b), chi) in itx.chis
        check(ccall((:pf_system_set_flory, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Float64),
                    sys.handle, a, b, chi))
    end
    return nothing
end

function add_interaction!(sys::HIPSystem, itx::EdwardsInteraction)
    for (m, u) in itx.uvals
        check(ccall((:pf_system_set_edwards, LIB), Cint, (Ptr{Cvoid}, Cint, Float64),
                    sys.handle, m, u))
    end
    return nothing
end

function set_omega!(sys::HIPSystem, mid::Integer, w::Array{Float64,3})
    @assert size(w) == sys.dims
    check(ccall((:pf_system_set_field, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                sys.handle, mid, w))
end

function getgrid(sys::HIPSystem, which::Integer, mid::Integer)
    out = Array{Float64,3}(undef, sys.dims)
    check(ccall((:pf_system_get_grid, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}),
                sys.handle, which, mid, out))
    return out
end

fields(sys::HIPSystem, mid) = getgrid(sys, 0, mid)
density(sys::HIPSystem, mid) = getgrid(sys, 1, mid)
residuals(sys::HIPSystem, mid) = getgrid(sys, 2, mid)

"""`fieldinit!(sys; seed, scale)` (src/fileio.jl:14-23), drawing with Julia's RNG."""
function fieldinit!(sys::HIPSystem; seed::Integer = -1, scale::Real = 0.1)
    if seed >= 0
        Random.seed!(seed)
    end
    # the draws follow the iteration order of the reference's Dict{Int} of fields, whose
    # keys are inserted in monomer order (src/system/system.jl:35,122)
    order = Dict{Int,Nothing}()
    for mid in sys.mids
        order[mid] = nothing
    end
    for mid in keys(order)
        set_omega!(sys, mid, scale * randn(sys.dims))
    end
    residuals!(sys)
end

density!(sys::HIPSystem) = check(ccall((:pf_density, LIB), Cint, (Ptr{Cvoid},), sys.handle))
residuals!(sys::HIPSystem) = check(ccall((:pf_residuals, LIB), Cint, (Ptr{Cvoid},), sys.handle))
validate(sys::HIPSystem) = check(ccall((:pf_validate, LIB), Cint, (Ptr{Cvoid},), sys.handle))

for (jl, c) in ((:field_error, :pf_field_error), (:free_energy, :pf_free_energy),
                (:free_energy_bulk, :pf_free_energy_bulk))
    @eval function $jl(sys::HIPSystem)
        v = Ref{Float64}(0.0)
        check(ccall(($(QuoteNode(c)), LIB), Cint, (Ptr{Cvoid}, Ref{Float64}), sys.handle, v))
        return v[]
    end
end

function scfstress(sys::HIPSystem)
    out = zeros(length(sys.cell.params))
    check(ccall((:pf_scfstress, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), sys.handle, out,
                length(out)))
    return out
end

mutable struct HIPUpdater
    handle::Ptr{Cvoid}
    kind::Symbol          # :anderson, :euler or :momentum
    params::NamedTuple    # the constructor's settings (batches read them)
end

function _updater(h::Ptr{Cvoid}, kind::Symbol, params::NamedTuple)
    u = HIPUpdater(h, kind, params)
    finalizer(x -> ccall((:pf_updater_destroy, LIB), Cint, (Ptr{Cvoid},), x.handle), u)
    return u
end

function AndersonUpdater(; nhist::Integer = 50, iter_euler::Integer = 1, lam_euler::Real = 0.01)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_updater_anderson, LIB), Cint, (Ref{Ptr{Cvoid}}, Cint, Cint, Float64), h,
                nhist, iter_euler, lam_euler))
    return _updater(h[], :anderson, (nhist = nhist, iter_euler = iter_euler, lam_euler = lam_euler))
end

"""`EulerUpdater(; method = :PECE, lam = 0.05)` (src/updaters/euler.jl:20-28)."""
function EulerUpdater(; method::Symbol = :PECE, lam::Real = 0.05)
    method in (:PEC, :PECE) ||
        throw(ArgumentError("Invalid `method` parameter for EulerUpdater. Options include: (:PEC, :PECE)"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_updater_euler, LIB), Cint, (Ref{Ptr{Cvoid}}, Cint, Float64), h,
                method == :PECE ? 1 : 0, lam))
    return _updater(h[], :euler, (method = method, lam = lam))
end

const PF_MOMENTUM = Dict(:MOM => Cint(0), :RMS => Cint(1), :ADAM => Cint(2), :NAG => Cint(3))

"""`MomentumUpdater(; method = :MOM, lam = 0.01, gamma = 0.9, beta1 = 0.9, beta2 = 0.999)`
(src/updaters/momentum.jl:30-41)."""
function MomentumUpdater(; method::Symbol = :MOM, lam::Real = 0.01, gamma::Real = 0.9,
                         beta1::Real = 0.9, beta2::Real = 0.999)
    haskey(PF_MOMENTUM, method) ||
        throw(ArgumentError("Invalid `method` parameter for MomentumUpdater."))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_updater_momentum, LIB), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Float64, Float64, Float64, Float64), h,
                PF_MOMENTUM[method], lam, gamma, beta1, beta2))
    return _updater(h[], :momentum,
                    (method = method, lam = lam, gamma = gamma, beta1 = beta1, beta2 = beta2))
end

setup!(u::HIPUpdater, sys::HIPSystem) =
    check(ccall((:pf_updater_setup, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), u.handle, sys.handle))
step!(u::HIPUpdater) = check(ccall((:pf_updater_step, LIB), Cint, (Ptr{Cvoid},), u.handle))

mutable struct HIPDomain
    handle::Ptr{Cvoid}
end

function DomainUpdater(; nper::Integer = 10, nskip::Integer = 200, lam::Real = 0.1,
                       pmax::Real = 1.0, rlxn::Real = 0.5)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pf_domain_create, LIB), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Cint, Float64, Float64, Float64), h, nper, nskip, lam,
                pmax, rlxn))
    d = HIPDomain(h[])
    finalizer(x -> ccall((:pf_domain_destroy, LIB), Cint, (Ptr{Cvoid},), x.handle), d)
    return d
end

struct PFOptions
    nsteps::Cint
    nsout::Cint
    nsave::Cint
    ftol::Cdouble
    stol::Cdouble
end

struct PFResults
    time::Cdouble
    nsteps::Cint
    converged::Cint
    ferr::Cdouble
    serr::Cdouble
    fhelm::Cdouble
end

"""
    scft!(sys, updater; domain = nothing, opts = SCFTOptions()) -> SCFTResults

`scft!` of src/simulate.jl:60-148; the whole loop runs inside the library.
"""
function scft!(sys::HIPSystem, updater::HIPUpdater; domain = nothing,
               opts::SCFTOptions = SCFTOptions())
    o = Ref(PFOptions(opts.nsteps, opts.nsout, opts.nsave, opts.ftol, opts.stol))
    r = Ref(PFResults(0, 0, 0, 0, 0, 0))
    dh = domain === nothing ? C_NULL : domain.handle
    if isempty(opts.savepath)
        check(ccall((:pf_scft, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{PFOptions}, Ref{PFResults}),
                    sys.handle, updater.handle, dh, o, r))
    else
        # savefields(sys, opts.savepath) every nsave steps (src/simulate.jl:124-128) and
        # after the final residuals!: the library calls back at those points
        job = (sys, opts.savepath)
        cb = @cfunction(_save_cb, Cvoid, (Any, Cint))
        check(ccall((:pf_scft_save, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{PFOptions}, Ref{PFResults},
                     Ptr{Cvoid}, Any),
                    sys.handle, updater.handle, dh, o, r, cb, job))
    end
    x = r[]
    return SCFTResults(x.time, x.nsteps, x.converged != 0, x.ferr, x.serr, x.fhelm)
end

# ---------------------------------------------------------------------------------------
# Batches (no counterpart in the reference): many systems advanced together.

mutable struct HIPBatch
    handle::Ptr{Cvoid}
    systems::Vector{HIPSystem}    # kept alive while the batch exists
end

function HIPBatch(systems::Vector{HIPSystem})
    h = Ref{Ptr{Cvoid}}(C_NULL)
    hs = Ptr{Cvoid}[s.handle for s in systems]
    check(ccall((:pf_batch_create, LIB), Cint, (Ref{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint), h, hs,
                length(hs)))
    b = HIPBatch(h[], systems)
    finalizer(x -> ccall((:pf_batch_destroy, LIB), Cint, (Ptr{Cvoid},), x.handle), b)
    return b
end

"""
    scft!(batch, nhist = 10; iter_euler = 1, lam_euler = 0.01, domain = nothing,
          opts = SCFTOptions(), mask = nothing) -> Vector{SCFTResults}
    scft!(batch, updater::HIPUpdater; domain = nothing, opts = SCFTOptions(),
          mask = nothing) -> Vector{SCFTResults}

`scft!(sys, AndersonUpdater(; nhist, iter_euler, lam_euler); domain, opts)` for every
system of the batch in lockstep (`pf_batch_scft`, or `pf_batch_scft_domain` with a
DomainUpdater, whose settings each system's own cell updater takes).  The second form
takes the settings of an AndersonUpdater, a MomentumUpdater (:MOM, :RMS, :ADAM, :NAG;
`pf_batch_scft_momentum`) or an EulerUpdater (:PEC, :PECE; `pf_batch_scft_euler`).  A
diverged system is reported with ferr = NaN instead of raising.
"""
function scft!(b::HIPBatch, u::HIPUpdater; domain = nothing,
               opts::SCFTOptions = SCFTOptions(), mask = nothing)
    p = u.params
    u.kind === :anderson && return scft!(b, p.nhist; iter_euler = p.iter_euler,
                                         lam_euler = p.lam_euler, domain = domain, opts = opts,
                                         mask = mask)
    r = Vector{PFResults}(undef, length(b.systems))
    m = mask === nothing ? C_NULL : Cint.(mask)
    dh = domain === nothing ? C_NULL : domain.handle
    if u.kind === :euler
        check(ccall((:pf_batch_scft_euler, LIB), Cint,
                    (Ptr{Cvoid}, Cint, Float64, Ptr{Cvoid}, Cint, Float64, Float64, Ptr{Cint},
                     Ptr{PFResults}),
                    b.handle, p.method === :PECE ? 1 : 0, p.lam, dh, opts.nsteps, opts.ftol,
                    opts.stol, m, r))
        return [SCFTResults(x.time, x.nsteps, x.converged != 0, x.ferr, x.serr, x.fhelm) for x in r]
    end
    check(ccall((:pf_batch_scft_momentum, LIB), Cint,
                (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Float64, Ptr{Cvoid}, Cint, Float64,
                 Float64, Ptr{Cint}, Ptr{PFResults}),
                b.handle, PF_MOMENTUM[p.method], p.lam, p.gamma, p.beta1, p.beta2, dh,
                opts.nsteps, opts.ftol, opts.stol, m, r))
    return [SCFTResults(x.time, x.nsteps, x.converged != 0, x.ferr, x.serr, x.fhelm) for x in r]
end

function scft!(b::HIPBatch, nhist::Integer = 10; iter_euler::Integer = 1,
               lam_euler::Real = 0.01, domain = nothing, opts::SCFTOptions = SCFTOptions(),
               mask = nothing)
    r = Vector{PFResults}(undef, length(b.systems))
    m = mask === nothing ? C_NULL : Cint.(mask)
    if domain === nothing
        check(ccall((:pf_batch_scft, LIB), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Float64, Cint, Float64, Ptr{Cint}, Ptr{PFResults}),
                    b.handle, nhist, iter_euler, lam_euler, opts.nsteps, opts.ftol, m, r))
    else
        check(ccall((:pf_batch_scft_domain, LIB), Cint,
                    (Ptr{Cvoid}, Cint, Cint, Float64, Ptr{Cvoid}, Cint, Float64, Float64,
                     Ptr{Cint}, Ptr{PFResults}),
                    b.handle, nhist, iter_euler, lam_euler, domain.handle, opts.nsteps,
                    opts.ftol, opts.stol, m, r))
    end
    return [SCFTResults(x.time, x.nsteps, x.converged != 0, x.ferr, x.serr, x.fhelm) for x in r]
end

# ---------------------------------------------------------------------------------------
# Reverse mode: d fhelm / d(w, cell, chi, u, phi, N_blk) (pf_fhelm_grad) and the
# ChainRulesCore rrule that exposes it to Zygote / Diffractor.

const PF_FHELM = Dict(:reference => Cint(0), :textbook => Cint(1))

"""
    fhelm_grad(sys; functional = :reference) -> (fhelm, dw::Dict{Int,Array{Float64,3}},
                                                 dcell, dchi::Matrix, du, dphi, dN)

Free energy and its gradient at the current fields (`pf_fhelm_grad`).  `:reference` is
`free_energy` as src/system/system.jl:402-433 computes it; `:textbook` uses
mu = log(phi / Q).  `dN` lists dF/dN_blk for every block of every chain species (contour
grid held fixed).
"""
function fhelm_grad(sys::HIPSystem; functional::Symbol = :reference)
    nm = length(sys.mids)
    n = prod(sys.dims)
    np = length(sys.cell.params)
    nb = sum(sp isa Multiblock ? length(sp.N_block) : 0 for sp in sys.species; init = 0)
    f = Ref{Float64}(0.0)
    dw = zeros(Float64, n * nm)
    dcell = zeros(np); dchi = zeros(nm, nm); du = zeros(nm)
    dphi = zeros(length(sys.species)); dN = zeros(max(nb, 1))
    check(ccall((:pf_fhelm_grad, LIB), Cint,
                (Ptr{Cvoid}, Cint, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
                sys.handle, PF_FHELM[functional], f, dw, dcell, np, dchi, du, dphi,
                nb > 0 ? dN : C_NULL, nb))
    # dw comes in slot order: ascending monomer ids
    dwd = Dict(mid => reshape(dw[(k - 1) * n + 1:k * n], sys.dims)
               for (k, mid) in enumerate(sort(sys.mids)))
    return f[], dwd, dcell, dchi, du, dphi, dN[1:nb]
end

"""
    density_vjp(sys, rho_bar::Dict, Q_bar::Vector) -> (dw, dcell, dN)

Vector-Jacobian product of `density!(sys)` (`pf_density_vjp`).
"""
function density_vjp(sys::HIPSystem, rho_bar::Dict, Q_bar::AbstractVector)
    nm = length(sys.mids)
    n = prod(sys.dims)
    rb = reduce(vcat, (vec(Float64.(rho_bar[mid])) for mid in sort(sys.mids)))
    dw = zeros(n * nm)
    dcell = zeros(length(sys.cell.params))
    nb = sum(sp isa Multiblock ? length(sp.N_block) : 0 for sp in sys.species; init = 0)
    dN = zeros(max(nb, 1))
    check(ccall((:pf_density_vjp, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint,
                 Ptr{Float64}, Cint),
                sys.handle, rb, Vector{Float64}(Q_bar), dw, dcell, length(dcell),
                nb > 0 ? dN : C_NULL, nb))
    return Dict(mid => reshape(dw[(k - 1) * n + 1:k * n], sys.dims)
                for (k, mid) in enumerate(sort(sys.mids))), dcell, dN[1:nb]
end

_save_cb(job, step::Cint) = (savefields(job[1], job[2]); nothing)

"""
    savefields(sys, fpath; field_type = :omega)

The reference's JSON field file (src/fileio.jl:77-100) from the device state.
"""
function savefields(sys::HIPSystem, fpath::AbstractString; field_type::Symbol = :omega)
    which = field_type == :rho ? 1 : 0
    d = Dict(:dim => sys.cell.dim, :grid => sys.dims, :mids => sort(sys.mids),
             :crystal => sys.cell.crystal, :cell_params => cellparams(sys),
             :field_type => field_type,
             :fields => Dict(mid => getgrid(sys, which, mid) for mid in sys.mids))
    open(fpath, "w") do f
        write(f, JSON.json(d, 4))
    end
    return nothing
end

function cellparams(sys::HIPSystem)
    p = zeros(length(sys.cell.params))
    check(ccall((:pf_system_cell, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint),
                sys.handle, p, length(p), 0))
    return p
end

"""`add_cell!(sys, cell)` (src/system/system.jl:102-106): the system takes `cell`."""
function add_cell!(sys::HIPSystem, cell::UnitCell)
    p = Vector{Float64}(cell.params)
    check(ccall((:pf_system_add_cell, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cstring, Ptr{Float64}, Cint),
                sys.handle, cell.dim, String(cell.crystal), p, length(p)))
    sys.cell = cell
    return nothing
end

"""
    set_checkpoint!(sys, K)

Checkpointed sweeps: a checkpoint every K contour slices (`-1`: ceil(sqrt(Ns + 1)), `0`:
full stacks).
"""
set_checkpoint!(sys::HIPSystem, K::Integer) =
    check(ccall((:pf_system_set_checkpoint, LIB), Cint, (Ptr{Cvoid}, Cint), sys.handle, K))

"""
    fhelm(sys, w::Dict, cellparams::Vector; functional = :reference)

The free energy as a function of the fields and the cell parameters: writes them into
`sys` and evaluates.  Differentiable through the rrule below.
"""
function fhelm(sys::HIPSystem, w::Dict, cellparams::AbstractVector; functional::Symbol = :reference)
    for (mid, grid) in w
        set_omega!(sys, mid, Array{Float64,3}(grid))
    end
    p = Vector{Float64}(cellparams)
    check(ccall((:pf_system_cell, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint),
                sys.handle, p, length(p), 1))
    return fhelm_grad(sys; functional = functional)[1]
end

# The rrule is defined when ChainRulesCore is loaded alongside this module.
function __init__()
    if isdefined(Main, :ChainRulesCore)
        CRC = getfield(Main, :ChainRulesCore)
        @eval function $(CRC).rrule(::typeof(fhelm), sys::HIPSystem, w::Dict,
                                    cellparams::AbstractVector; functional::Symbol = :reference)
            for (mid, grid) in w
                set_omega!(sys, mid, Array{Float64,3}(grid))
            end
            p = Vector{Float64}(cellparams)
            check(ccall((:pf_system_cell, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint),
                        sys.handle, p, length(p), 1))
            F, dw, dcell = fhelm_grad(sys; functional = functional)
            pullback(dF) = ($(CRC).NoTangent(), $(CRC).NoTangent(),
                            Dict(mid => dF .* g for (mid, g) in dw), dF .* dcell)
            return F, pullback
        end
    end
end

using Random
import JSON

end # module

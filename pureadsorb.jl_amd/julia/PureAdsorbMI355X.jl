"""
    PureAdsorbMI355X

The MI355X backend of PureAdsorb.jl's Widom and move-energy path: Julia bindings over the C ABI
of `libpureadsorb_amd.so` (`include/pureadsorb_amd.h`). Indices cross as 0-based; 3x3 matrices
and `Vector{SVector{3,Float64}}` cross as their own column-major memory.

Entry points:
- `widom(batch::FrameworkBatch, guest; backend = MI355X(), T, ninsert, seed, chunk, nblocks, run)`
  through the `widom_external` hook (two lines a maintainer adds to `src/widom.jl`, see
  `INTEGRATION.md`), or directly as `widom_mi355x(batch, guest; ...)`: the reference's own
  `FrameworkBatch` is taken over as it stands (`pa_batch_from_tables`);
- `MI355XBatch(fws, ff, guest, ewald; fullk)` builds the tables on the library side
  (`pa_batch_create`) and serves `widom`, `widom_grad` (with a ChainRulesCore `rrule` over
  ln K_H) and the chains (`MI355XState`, `mc_step!`, `run_gcmc!`, `run_nvt!`).

Julia is not installed where this library is built, so this file has not been executed there;
the same entry points are exercised through Python's ctypes by the library's test suite.
"""
module PureAdsorbMI355X

using PureAdsorb
using PureAdsorb: FrameworkBatch, Framework, ForceField, Guest, EwaldParams, WidomResult, GCMCResult, NVTResult
# methods on this module's own batch and state types join the reference's functions
import PureAdsorb: widom, mc_step!, run_gcmc!, run_nvt!
using StaticArrays
using Unitful: Unitful, ustrip, @u_str
using ChainRulesCore

export MI355X, MI355XBatch, MI355XState, widom_mi355x, widom_grad

const LIB = get(ENV, "PUREADSORB_AMD_LIB", "libpureadsorb_amd.so")
const PREC_F64, PREC_MIXED = UInt32(0), UInt32(1)
const BATCH_FULLK = UInt32(1)

"""Backend selector: `widom(batch, guest; backend = MI355X(), ...)`."""
struct MI355X end

_kelvin(T::Real) = Float64(T)
_kelvin(T::Unitful.Quantity) = Float64(ustrip(u"K", T))
_pascal(p::Real) = Float64(p)
_pascal(p::Unitful.Quantity) = Float64(ustrip(u"Pa", p))
_precision(p::Symbol) = p === :f64 ? PREC_F64 : p === :mixed ? PREC_MIXED :
                        throw(ArgumentError("precision=$p: use :mixed or :f64"))

# ---- C structures (include/pureadsorb_amd.h), field for field
struct PaFramework
    cell::NTuple{9, Float64}
    natoms::Int32
    frac::Ptr{Float64}
    types::Ptr{Int32}
    charges::Ptr{Float64}
    replication::NTuple{3, Int32}
end
struct PaForcefield
    ntypes::Int32
    sigma::Ptr{Float64}
    epsilon::Ptr{Float64}
    cutoff::Float64
    tail::Int32
end
struct PaGuest
    nsites::Int32
    sites::Ptr{Float64}
    types::Ptr{Int32}
    charges::Ptr{Float64}
end
struct PaEwald
    cutoff::Float64
    precision::Float64
end
struct PaWidomResult
    mu_ex::Float64; mu_ex_err::Float64; K_H::Float64; K_H_err::Float64
    q_st::Float64; q_st_err::Float64; nsamples::Int64; nblocks::Int64; ln_K_H::Float64
end
struct PaBatchTables
    nframeworks::Int32
    nsys::Int32
    framework_of::Ptr{Int32}
    atom_offsets::Ptr{Int32}
    positions::Ptr{Float64}
    types::Ptr{Int32}
    charges::Ptr{Float64}
    cells::Ptr{Float64}
    alphas::Ptr{Float64}
    k_offsets::Ptr{Int32}
    ks::Ptr{Float64}
    kprefactor::Ptr{Float64}
    Shost::Ptr{Float64}
    constant_offset::Ptr{Float64}
    self_term_halfrange::Ptr{Float64}
    ntypes::Int32
    sigma::Ptr{Float64}
    epsilon::Ptr{Float64}
    compact_to_orig::Ptr{Int32}
    bs::Ptr{Float64}
    kmin::Ptr{Float64}
    cutoff::Float64
    ewald_cutoff::Float64
    fullk::Int32
end
struct PaGcmcParams
    T::Float64; n_warmup::Int64; n_production::Int64; n_audit::Int64; min_cycle_length::Int64
    fugacity_pa::Ptr{Float64}; step_trans::Ptr{Float64}; step_rot::Ptr{Float64}
    p_translation::Float64; p_rotation::Float64; p_reinsertion::Float64; p_exchange::Float64
    nblocks::Int32
end
struct PaGcmcResult
    loading::Float64; loading_err::Float64; energy::Float64; energy_err::Float64
    q_st::Float64; q_st_err::Float64; corr_UN::Float64; max_occupancy::Int32; capacity::Int32
    ncycles::Int64; accepted::NTuple{5, Int64}; attempted::NTuple{5, Int64}
end
struct PaNvtParams
    T::Float64; n_warmup::Int64; n_production::Int64; n_widom_per_cycle::Int64; n_audit::Int64
    min_cycle_length::Int64; step_trans::Ptr{Float64}; step_rot::Ptr{Float64}; widom_seed::UInt64
    nblocks::Int32
end
struct PaNvtResult
    energy::Float64; energy_err::Float64; mu_ex::Float64; mu_ex_err::Float64
    K_H::Float64; K_H_err::Float64; q_st::Float64; q_st_err::Float64
    acceptance::NTuple{3, Float64}; ncycles::Int64
end

function check(st::Integer)
    st == 0 && return nothing
    buf = Vector{UInt8}(undef, 4096)
    ccall((:pa_last_error, LIB), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf))
    msg = unsafe_string(pointer(buf))
    st == 2 && throw(DimensionMismatch(msg))
    st == 1 && throw(ArgumentError(msg))
    return error("libpureadsorb_amd status $st: $msg")
end

"""The guest as the C ABI takes it; `buffers` keeps the arrays alive (GC.@preserve it)."""
struct GuestArgs
    buffers::Tuple{Vector{Float64}, Vector{Int32}, Vector{Float64}}
end
function GuestArgs(g::Guest{Float64})
    sites = collect(reinterpret(Float64, collect(g.sites)))
    types = Int32.(collect(g.types) .- 1)
    charges = collect(Float64, g.charges)
    return GuestArgs((sites, types, charges))
end
pa_guest(ga::GuestArgs) = PaGuest(Int32(length(ga.buffers[3])), pointer(ga.buffers[1]), pointer(ga.buffers[2]),
                                  pointer(ga.buffers[3]))

_result(r::PaWidomResult) = WidomResult(r.mu_ex, r.mu_ex_err, r.K_H, r.K_H_err, r.q_st, r.q_st_err, Int(r.nsamples),
                                        Int(r.nblocks))

# ---- batches
"""A batch's device tables, owned by the library; freed by a finalizer."""
mutable struct MI355XBatch
    handle::Ptr{Cvoid}
    nsys::Int
    natoms::Vector{Int}   # per system: atoms of its framework (charge-gradient layout)
    ntypes::Int           # force-field types (LJ-gradient layout); 0 for a batch taken from tables
    function MI355XBatch(h::Ptr{Cvoid}, nsys::Integer, natoms::Vector{Int}, ntypes::Integer)
        b = new(h, Int(nsys), natoms, Int(ntypes))
        finalizer(x -> ccall((:pa_batch_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.handle), b)
        return b
    end
end

"""`FrameworkBatch(fws, ff, guest, ewald; fullk)` (src/batch.jl:237) built by the library; atom
types map through `typeindex(ff, symbol * "_")` as the reference does (src/batch.jl:309)."""
function MI355XBatch(fws::AbstractVector{<:Framework{Float64}}, ff::ForceField{Float64}, guest::Guest{Float64},
                     ewald::EwaldParams{Float64}; fullk::Bool = false)
    keep = Any[]
    pfw = map(fws) do fw
        frac = collect(reinterpret(Float64, fw.frac))
        types = Int32[PureAdsorb.typeindex(ff, s * "_") - 1 for s in fw.symbols]
        q = collect(Float64, fw.charges)
        push!(keep, (frac, types, q))
        PaFramework(Tuple(fw.cell), Int32(length(q)), pointer(frac), pointer(types), pointer(q), Int32.(fw.replication))
    end
    sig, eps = collect(Float64, ff.sigma), collect(Float64, ff.epsilon)   # column-major already
    ga = GuestArgs(guest)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep sig eps ga begin
        pff = Ref(PaForcefield(Int32(size(sig, 1)), pointer(sig), pointer(eps), ff.cutoff, Int32(ff.tail)))
        pg = Ref(pa_guest(ga))
        pe = Ref(PaEwald(ewald.cutoff, ewald.precision))
        check(ccall((:pa_batch_create, LIB), Cint,
                    (Ptr{PaFramework}, Int32, Ref{PaForcefield}, Ref{PaGuest}, Ref{PaEwald}, UInt32, Ref{Ptr{Cvoid}}),
                    pfw, Int32(length(pfw)), pff, pg, pe, fullk ? BATCH_FULLK : UInt32(0), h))
    end
    return MI355XBatch(h[], length(fws), [length(fw.charges) for fw in fws], size(sig, 1))
end

"""The reference's own `FrameworkBatch` (src/batch.jl:72-104) taken over by the library
(`pa_batch_from_tables`): its tables are copied as they are, 1-based indices made 0-based."""
function MI355XBatch(batch::FrameworkBatch{Float64}, guest::Guest{Float64})
    fo = Int32.(collect(batch.framework_of) .- 1)
    ao = Int32.(collect(batch.atom_offsets))
    pos = collect(reinterpret(Float64, collect(batch.positions)))
    types = Int32.(collect(batch.types) .- 1)
    q = collect(Float64, batch.charges)
    cells = collect(reinterpret(Float64, collect(batch.cells)))
    alphas = collect(Float64, batch.alphas)
    ko = Int32.(collect(batch.k_offsets))
    ks = collect(reinterpret(Float64, collect(batch.ks)))
    kp = collect(Float64, batch.kprefactor)
    sh = collect(reinterpret(Float64, collect(ComplexF64, batch.Shost)))
    co = collect(Float64, batch.constant_offset)
    hr = collect(Float64, batch.self_term_halfrange)
    sig, eps = collect(Float64, batch.sigma), collect(Float64, batch.epsilon)
    c2o = Int32.(collect(batch.compact_to_orig) .- 1)
    bs = collect(Float64, batch.bs)
    kmin = collect(Float64, batch.kmin)
    ga = GuestArgs(guest)
    nfw = length(alphas)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve fo ao pos types q cells alphas ko ks kp sh co hr sig eps c2o bs kmin ga begin
        t = Ref(PaBatchTables(Int32(nfw), Int32(batch.nsys), pointer(fo), pointer(ao), pointer(pos), pointer(types),
                              pointer(q), pointer(cells), pointer(alphas), pointer(ko), pointer(ks), pointer(kp),
                              pointer(sh), pointer(co), pointer(hr), Int32(size(sig, 1)), pointer(sig), pointer(eps),
                              pointer(c2o), pointer(bs), pointer(kmin), Float64(batch.cutoff),
                              Float64(batch.ewald_cutoff), Int32(batch.fullk)))
        pg = Ref(pa_guest(ga))
        check(ccall((:pa_batch_from_tables, LIB), Cint, (Ref{PaBatchTables}, Ref{PaGuest}, Ref{Ptr{Cvoid}}), t, pg, h))
    end
    natoms = [Int(ao[f + 1] - ao[f]) for f in (fo .+ 1)]
    return MI355XBatch(h[], batch.nsys, natoms, 0)
end

# ---- Widom
"""`widom` (src/widom.jl:255-301) on the device. `run = nothing` selects `default_run`; `chunk`
is the insertions per energy launch (the library picks one when `nothing`); `precision` is
`:mixed` (FP32 pair and k-space terms, FP64 sums) or `:f64`; `reject = false` is the
reference's `widom_singlephase`."""
function widom(b::MI355XBatch, guest::Guest{Float64}; T, ninsert::Integer, seed::Integer = 0, chunk = nothing,
               nblocks::Integer = 10, run = nothing, precision::Symbol = :mixed, reject::Bool = true,
               no_alloc::Bool = false)
    nblocks >= 2 || throw(ArgumentError("nblocks=$nblocks: at least two blocks are needed for a standard error"))
    isnothing(chunk) || chunk >= 1 || throw(ArgumentError("chunk=$chunk must be ≥ 1"))
    isnothing(run) || run >= 1 || throw(ArgumentError("run=$run must be ≥ 1"))
    ga = GuestArgs(guest)
    out = Vector{PaWidomResult}(undef, b.nsys)
    prec = _precision(precision) | (reject ? UInt32(0) : UInt32(0x100))
    GC.@preserve ga begin
        pg = Ref(pa_guest(ga))
        st = if no_alloc
            ccall((:pa_widom_noalloc, LIB), Cint,
                  (Ptr{Cvoid}, Ref{PaGuest}, Float64, Int64, UInt64, Int32, Int32, Int64, UInt32, Ptr{PaWidomResult}),
                  b.handle, pg, _kelvin(T), Int64(ninsert), seed % UInt64, Int32(nblocks), Int32(something(run, 0)),
                  Int64(something(chunk, 0)), prec, out)
        else
            ccall((:pa_widom, LIB), Cint,
                  (Ptr{Cvoid}, Ref{PaGuest}, Float64, Int64, UInt64, Int32, Int32, Int64, UInt32, Ptr{PaWidomResult}),
                  b.handle, pg, _kelvin(T), Int64(ninsert), seed % UInt64, Int32(nblocks), Int32(something(run, 0)),
                  Int64(something(chunk, 0)), prec, out)
        end
        check(st)
    end
    return [_result(r) for r in out]
end

"""`widom(batch::FrameworkBatch, guest; ...)` on the MI355X: the reference's batch taken over as it
stands, then `widom` on the device."""
widom_mi355x(batch::FrameworkBatch{Float64}, guest::Guest{Float64}; kwargs...) =
    widom(MI355XBatch(batch, guest), guest; kwargs...)

# The backend hook: with `widom_external` declared in PureAdsorb and called from `_widom` for a
# backend that is not a KernelAbstractions backend (INTEGRATION.md), the reference's own call
# `widom(batch, guest; backend = MI355X(), T, ninsert, ...)` lands here.
if isdefined(PureAdsorb, :widom_external)
    function PureAdsorb.widom_external(::MI355X, batch::FrameworkBatch{Float64}, guest::Guest{Float64}; reject::Bool = true,
                                       T, ninsert::Integer, seed::Integer = 0, chunk::Integer = 2^16,
                                       nblocks::Integer = 10, run = nothing)
        return widom_mi355x(batch, guest; reject, T, ninsert, seed, chunk, nblocks, run)
    end
end

# ---- gradients (no counterpart in the reference)
"""Gradients of ln K_H and q_st per system: per force-field type ε and σ (Lorentz-Berthelot) and per
framework-atom charge, from the same insertions as `widom`. Returns `(results, lj, dq)` with
`lj = (dlnkh_deps, dlnkh_dsig, dqst_deps, dqst_dsig)`, each `ntypes x nsys`, and
`dq = (dlnkh_dq, dqst_dq)`, each a vector per system."""
function widom_grad(b::MI355XBatch, guest::Guest{Float64}; T, ninsert::Integer, seed::Integer = 0,
                    nblocks::Integer = 10, run = nothing, precision::Symbol = :mixed)
    b.ntypes > 0 || throw(ArgumentError("widom_grad needs a batch built from Frameworks and a ForceField"))
    ga = GuestArgs(guest)
    out = Vector{PaWidomResult}(undef, b.nsys)
    lj = [zeros(b.ntypes * b.nsys) for _ in 1:4]
    dq = [zeros(sum(b.natoms)) for _ in 1:2]
    GC.@preserve ga begin
        pg = Ref(pa_guest(ga))
        check(ccall((:pa_widom_grad, LIB), Cint,
                    (Ptr{Cvoid}, Ref{PaGuest}, Float64, Int64, UInt64, Int32, Int32, UInt32, Ptr{PaWidomResult},
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    b.handle, pg, _kelvin(T), Int64(ninsert), seed % UInt64, Int32(nblocks), Int32(something(run, 0)),
                    _precision(precision), out, lj[1], lj[2], lj[3], lj[4], dq[1], dq[2]))
    end
    offs = cumsum(vcat(0, b.natoms))
    split(v) = [v[offs[s] + 1:offs[s + 1]] for s in 1:b.nsys]
    return [_result(r) for r in out], Tuple(reshape(x, b.ntypes, b.nsys) for x in lj), (split(dq[1]), split(dq[2]))
end

"""ln K_H of every framework as a function of the per-type ε, σ and the per-framework charge
vectors, through the device; `lnKH_parts` carries what does not vary (names, frameworks, guest,
cutoffs, Ewald parameters, Widom settings)."""
struct LnKHProblem
    names::Vector{String}
    frameworks::Vector{Framework{Float64}}
    guest::Guest{Float64}
    cutoff::Float64
    ewald::EwaldParams{Float64}
    widom_kwargs::NamedTuple
end

function lnKH(p::LnKHProblem, eps::AbstractVector{Float64}, sig::AbstractVector{Float64},
              charges::AbstractVector{<:AbstractVector{Float64}})
    ff = ForceField(p.names, collect(sig), collect(eps); cutoff = p.cutoff, tail = false)
    fws = [Framework{Float64}(fw.cell, fw.frac, fw.labels, fw.symbols, collect(q), fw.replication)
           for (fw, q) in zip(p.frameworks, charges)]
    b = MI355XBatch(fws, ff, p.guest, p.ewald)
    res, lj, dq = widom_grad(b, p.guest; p.widom_kwargs...)
    return [log(r.K_H) for r in res], (lj, dq)
end

function ChainRulesCore.rrule(::typeof(lnKH), p::LnKHProblem, eps, sig, charges)
    y, (lj, dq) = lnKH(p, eps, sig, charges)
    function lnKH_pullback(ȳ)
        ȳv = collect(Float64, unthunk(ȳ))
        deps = lj[1] * ȳv                       # ntypes x nsys times nsys
        dsig = lj[2] * ȳv
        dch = [dq[1][s] .* ȳv[s] for s in eachindex(ȳv)]
        return NoTangent(), NoTangent(), deps, dsig, dch
    end
    return y, lnKH_pullback
end

# ---- chains: SystemState and the move path
"""`SystemState(batch, guest, ncounts; T, seed, capacities)` (src/state.jl:328-446) on the device; the
batch must be built with `fullk = true`."""
mutable struct MI355XState
    handle::Ptr{Cvoid}
    nsys::Int
    batch::MI355XBatch
    function MI355XState(b::MI355XBatch, guest::Guest{Float64}, ncounts::AbstractVector{<:Integer}; T, seed::Integer = 0,
                         capacities = nothing)
        length(ncounts) == b.nsys || throw(DimensionMismatch("ncounts must have one entry per system (nsys=$(b.nsys))"))
        nc = Int32.(collect(ncounts))
        cap = isnothing(capacities) ? nothing : Int32.(collect(capacities))
        ga = GuestArgs(guest)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve nc cap ga begin
            pg = Ref(pa_guest(ga))
            check(ccall((:pa_state_create, LIB), Cint,
                        (Ptr{Cvoid}, Ref{PaGuest}, Ptr{Int32}, Ptr{Int32}, Float64, UInt64, Ref{Ptr{Cvoid}}),
                        b.handle, pg, nc, isnothing(cap) ? Ptr{Int32}(C_NULL) : pointer(cap), _kelvin(T),
                        seed % UInt64, h))
        end
        s = new(h[], b.nsys, b)
        finalizer(x -> ccall((:pa_state_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.handle), s)
        return s
    end
end

_pervec(x, n, name) = (v = x isa Number ? fill(Float64(x), n) : Float64.(collect(x));
                       length(v) == n || throw(DimensionMismatch("$name must have one entry per system (nsys=$n)")); v)

"""`mc_step!` (src/moves.jl:577-645): one attempt of `movetype` (1 translation, 2 rotation,
3 reinsertion) on every chain."""
function mc_step!(st::MI355XState, movetype::Integer, step_trans, step_rot; T)
    a, r = _pervec(step_trans, st.nsys, "step_trans"), _pervec(step_rot, st.nsys, "step_rot")
    check(ccall((:pa_mc_step, LIB), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Float64),
                st.handle, Int32(movetype), a, r, _kelvin(T)))
    return nothing
end

"""`run_gcmc!` (src/gcmc.jl:161-337) on the device: an exchange (fair insertion/deletion coin) with
probability `exchange_prob`, else translation, rotation or reinsertion with equal probability,
each drawn from the chain's own counter-based stream."""
function run_gcmc!(st::MI355XState; T, n_warmup::Integer, n_production::Integer, n_audit::Integer, step_trans,
                   step_rot, fugacity, exchange_prob::Real = 0.5, min_cycle_length::Integer = 1, nblocks::Integer = 10)
    n = st.nsys
    f = [_pascal(x) for x in (fugacity isa AbstractVector ? fugacity : fill(fugacity, n))]
    length(f) == n || throw(DimensionMismatch("fugacity must have one entry per system (nsys=$n)"))
    a, r = _pervec(step_trans, n, "step_trans"), _pervec(step_rot, n, "step_rot")
    pm = (1.0 - Float64(exchange_prob)) / 3
    out = Vector{PaGcmcResult}(undef, n)
    GC.@preserve f a r begin
        p = Ref(PaGcmcParams(_kelvin(T), n_warmup, n_production, n_audit, min_cycle_length, pointer(f), pointer(a),
                             pointer(r), pm, pm, pm, Float64(exchange_prob), Int32(nblocks)))
        check(ccall((:pa_gcmc_run, LIB), Cint, (Ptr{Cvoid}, Ref{PaGcmcParams}, Ptr{PaGcmcResult}), st.handle, p, out))
    end
    return [GCMCResult{Float64}(x.loading, x.loading_err, x.energy, x.energy_err, x.q_st, x.q_st_err, x.corr_UN,
                                Int(x.max_occupancy), Int(x.capacity), Int(x.ncycles)) for x in out]
end

"""`run_nvt!` (src/nvt.jl:192-358) on the device, test insertions from `widom_seed`'s stream."""
function run_nvt!(st::MI355XState, guest::Guest{Float64}; T, n_warmup::Integer, n_production::Integer,
                  n_widom_per_cycle::Integer, n_audit::Integer, step_trans, step_rot, min_cycle_length::Integer = 1,
                  widom_seed::Integer = 0, nblocks::Integer = 10)
    n = st.nsys
    a, r = _pervec(step_trans, n, "step_trans"), _pervec(step_rot, n, "step_rot")
    ga = GuestArgs(guest)
    out = Vector{PaNvtResult}(undef, n)
    GC.@preserve a r ga begin
        p = Ref(PaNvtParams(_kelvin(T), n_warmup, n_production, n_widom_per_cycle, n_audit, min_cycle_length,
                            pointer(a), pointer(r), widom_seed % UInt64, Int32(nblocks)))
        pg = Ref(pa_guest(ga))
        check(ccall((:pa_nvt_run, LIB), Cint, (Ptr{Cvoid}, Ref{PaGuest}, Ref{PaNvtParams}, Ptr{PaNvtResult}),
                    st.handle, pg, p, out))
    end
    return [NVTResult{Float64}(x.energy, x.energy_err, x.mu_ex, x.mu_ex_err, x.K_H, x.K_H_err, x.q_st, x.q_st_err,
                               SVector(x.acceptance), Int(x.ncycles)) for x in out]
end

end # module
